// acouav main (src/Setup.cpp:73-129): parse, read TSPLIB, distances, colony,
// header, run.
#include <iomanip>
#include <iostream>
#include <memory>

#include "AntSystem.h"
#include "RankBasedAntSystem.h"
#include "acouav.h"

int main(int argc, char* argv[]) {
    std::cout << std::setprecision(8);
    Commandline cl;
    if (!cl.parse(argc, argv)) {
        cl.printHelp();
        return -1;
    }
    TSPReader reader;
    if (!reader.read(cl.tsplibfile)) {
        std::cout << "\n" << reader.error << "\nUnable to read tsplib file [" << cl.tsplibfile << "]\n";
        cl.printHelp();
        return -1;
    }
    Stopwatch sw;
    sw.start();
    std::vector<int> dist = euc2dMatrix(reader.x, reader.y);
    const float calcdisttime = sw.elapsed();

    std::unique_ptr<Colony> colony;
    if (cl.rankBased) {
        auto* rb = new RankBasedAntSystem(dist.data(), reader.nodes, cl.numAnts);
        rb->setAlpha(cl.alpha);
        rb->setBeta(cl.beta);
        rb->setRho(cl.rho);
        rb->setW(cl.w);
        rb->setDorigoInitialPheromone(cl.dorigoTau0);
        colony.reset(rb);
    } else if (cl.aging) {
        colony.reset(new AntSystem(dist.data(), reader.nodes, cl.numAnts, cl.alpha, cl.beta, cl.rho,
                                   cl.minRetirementAge, cl.maxRetirementAge, cl.localPeformance,
                                   cl.globalPerformance, cl.minPopulationLimit, cl.maxPopulationLimit));
    } else {
        colony.reset(new AntSystem(dist.data(), reader.nodes, cl.numAnts, cl.alpha, cl.beta, cl.rho));
    }
    if (cl.aging) {
        colony->setAging(true);
        colony->setMinRetirementAge(cl.minRetirementAge);
        colony->setMaxRetirementAge(cl.maxRetirementAge);
        colony->setLocalPeformance(cl.localPeformance);
        colony->setGlobalPerformance(cl.globalPerformance);
        colony->setMinPopulationLimit(cl.minPopulationLimit);
        colony->setMaxPopulationLimit(cl.maxPopulationLimit);
    }
    colony->setBroker(cl.broker);
    colony->setCandidateListLength(cl.candidates);
    colony->setSeed(cl.seed);
    Writer::VERBOSE = cl.verbose;
    {
        Writer w(cl.outfile);
        w.writeHeader("GPU", cl.alpha, cl.beta, cl.rho, cl.numAnts, colony->antSystemName(), reader.name,
                      calcdisttime, cl.aging);
        w.close();
    }
    TSP tsp(cl.maxTime, cl.maxIter, cl.maxReps, cl.optimum);
    tsp.execute(reader.name.c_str(), *colony, cl.outfile.c_str(), calcdisttime, cl.aging);
    return 0;
}
