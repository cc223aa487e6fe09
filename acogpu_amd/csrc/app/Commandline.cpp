// acouav flags: the reference set (src/Commandline.cpp:14-231) with its
// defaults (inc/default.h:8-39), plus the selections this build makes at run
// time instead of build time.
#include <cstdio>
#include <cstdlib>
#include <iostream>

#include "acouav.h"

static const char* kHelp =
    "  -a, --alpha #            weight of the pheromone term (tau^alpha)\n"
    "  -b, --beta #             weight of the distance heuristic (eta^beta)\n"
    "  -r, --rho #              evaporation rate per iteration, 0..1\n"
    "  -s, --tours #            iteration limit\n"
    "  -t, --time #             time limit in ms (summed phase times)\n"
    "  -i, --tsplibfile FILE    instance in TSPLIB format (EUC_2D)\n"
    "  -p, --outfile FILE       per-iteration statistics are written here\n"
    "  -o, --optimum #          stop once an iteration's best tour is this short\n"
    "  -m, --ants #             colony size\n"
    "  -c, --repetitions #      repetition limit\n"
    "  -j, --minage #           aging: below this age after convergence an ant is young\n"
    "  -k, --maxage #           aging: ants of this age retire (one per iteration)\n"
    "  -l, --localperformance # aging: accepted, not used by the rules\n"
    "  -g, --globalperformance # aging: recruit after this many iterations without improvement\n"
    "  -x, --minpopulation #    aging: recruit while fewer ants are valid\n"
    "  -y, --maxpopulation #    aging: accepted, not used by the rules\n"
    "  -v, --verbose            echo the statistics to the console\n"
    "  -h, --help           display this help text and exit\n"
    " MI355X build selections:\n"
    "      --aging / --noaging       ant aging variant (default from the program name)\n"
    "      --broker / --nobroker     broker variant (no behaviour)\n"
    "      --rank-based             rank-based Ant System\n"
    "      --w #                    ranked ants of the rank-based system (default 5)\n"
    "      --dorigo-tau0            tau0 = 0.5 w (w-1) / (rho C_nn) for the rank-based system\n"
    "      --candidates #           candidate list length k (1..64); 0 = every city (default)\n"
    "      --seed #                 random stream key (default 12345678)\n"
    "      --host-distances         hand the distance matrix to the colony in host memory\n"
    "                               (default: device memory, as the reference's GPU build)\n";

Commandline::Commandline()
    : numAnts(10), maxTime(0x7fffffff), maxIter(10), maxReps(0x7fffffff), optimum(-1), alpha(1.0f), beta(5.0f),
      rho(0.1f), minRetirementAge(30), maxRetirementAge(50), localPeformance(10), globalPerformance(10),
      minPopulationLimit(40), maxPopulationLimit(100) {}

bool Commandline::parse(int argc, char* argv[]) {
    // The build variant comes from the program name, as the reference's
    // acouav_<platform>_<aging>_<broker> binaries (release/makefile:52-59).
    if (argc > 0) {
        std::string prog = argv[0];
        prog = prog.substr(prog.find_last_of('/') == std::string::npos ? 0 : prog.find_last_of('/') + 1);
        if (prog.find("_noaging") != std::string::npos)
            aging = false;
        else if (prog.find("_aging") != std::string::npos)
            aging = true;
        if (prog.find("_nobroker") != std::string::npos)
            broker = false;
        else if (prog.find("_broker") != std::string::npos)
            broker = true;
    }
    auto next = [&](int& i) -> const char* {
        if (i + 1 >= argc) {
            std::fprintf(stderr, "missing value after %s\n", argv[i]);
            std::exit(EXIT_FAILURE);
        }
        return argv[++i];
    };
    for (int i = 1; i < argc; ++i) {
        const std::string o = argv[i];
        if (o == "-a" || o == "--alpha") alpha = (float)std::atof(next(i));
        else if (o == "-b" || o == "--beta") beta = (float)std::atof(next(i));
        else if (o == "-r" || o == "--rho") rho = (float)std::atof(next(i));
        else if (o == "-i" || o == "--tsplibfile") tsplibfile = next(i);
        else if (o == "-p" || o == "--outfile") outfile = next(i);
        else if (o == "-m" || o == "--ants") numAnts = std::atoi(next(i));
        else if (o == "-t" || o == "--time") maxTime = std::atoi(next(i));
        else if (o == "-s" || o == "--tours") maxIter = std::atoi(next(i));
        else if (o == "-o" || o == "--optimum") optimum = std::atoi(next(i));
        else if (o == "-c" || o == "--repetitions") maxReps = std::atoi(next(i));
        else if (o == "-v" || o == "--verbose") verbose = true;
        else if (o == "-h" || o == "--help") return false;
        else if (o == "-j" || o == "--minage") minRetirementAge = std::atoi(next(i));
        else if (o == "-k" || o == "--maxage") maxRetirementAge = std::atoi(next(i));
        else if (o == "-l" || o == "--localperformance") localPeformance = std::atoi(next(i));
        else if (o == "-g" || o == "--globalperformance") globalPerformance = std::atoi(next(i));
        else if (o == "-x" || o == "--minpopulation") minPopulationLimit = std::atoi(next(i));
        else if (o == "-y" || o == "--maxpopulation") maxPopulationLimit = std::atoi(next(i));
        else if (o == "--aging") aging = true;
        else if (o == "--noaging") aging = false;
        else if (o == "--broker") broker = true;
        else if (o == "--nobroker") broker = false;
        else if (o == "--rank-based") rankBased = true;
        else if (o == "--w") w = std::atoi(next(i));
        else if (o == "--dorigo-tau0") dorigoTau0 = true;
        else if (o == "--candidates") candidates = std::atoi(next(i));
        else if (o == "--seed") seed = std::strtoull(next(i), nullptr, 10);
        else if (o == "--host-distances") hostDistances = true;
        // unknown arguments are ignored, as in the reference
    }
    return true;
}

void Commandline::printHelp() const { std::cout << "Usage: [OPTION]... [ARGUMENT]...\n\n Options:\n" << kHelp; }
