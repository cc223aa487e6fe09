// acouav main (src/Setup.cpp:73-129): parse, read TSPLIB, distances, colony,
// header, run.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
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
    // The reference's GPU build hands the colony a device copy of the matrix
    // and frees it after the run (src/Setup.cpp:60-66, 120-124).
    int* distances = dist.data();
    int* d_dist = nullptr;
    if (!cl.hostDistances) {
        const size_t bytes = dist.size() * sizeof(int);
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&d_dist), bytes);
        if (e == hipSuccess) e = hipMemcpy(d_dist, dist.data(), bytes, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            std::fprintf(stderr, "device copy of the distance matrix: %s\n", hipGetErrorString(e));
            return EXIT_FAILURE;
        }
        distances = d_dist;
    }
    const float calcdisttime = sw.elapsed();

    std::unique_ptr<Colony> colony;
    if (cl.rankBased) {
        auto* rb = new RankBasedAntSystem(distances, reader.nodes, cl.numAnts);
        rb->setAlpha(cl.alpha);
        rb->setBeta(cl.beta);
        rb->setRho(cl.rho);
        rb->setW(cl.w);
        rb->setDorigoInitialPheromone(cl.dorigoTau0);
        colony.reset(rb);
    } else if (cl.aging) {
        colony.reset(new AntSystem(distances, reader.nodes, cl.numAnts, cl.alpha, cl.beta, cl.rho,
                                   cl.minRetirementAge, cl.maxRetirementAge, cl.localPeformance,
                                   cl.globalPerformance, cl.minPopulationLimit, cl.maxPopulationLimit));
    } else {
        colony.reset(new AntSystem(distances, reader.nodes, cl.numAnts, cl.alpha, cl.beta, cl.rho));
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
    colony.reset();
    if (d_dist) (void)hipFree(d_dist);
    return 0;
}
