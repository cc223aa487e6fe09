// Colony: the drop-in C++ surface of the reference's ACO engine
// (inc/Colony.h:24-155), implemented over the C ABI of include/acogpu.h.
//
// The public member functions keep the reference's names, arguments and
// meaning.  Differences a caller can observe:
//   * settings are taken when initialize() creates the device context, so
//     setters called after initialize() do not change a running colony;
//   * the aging variant is a run-time switch (setAging) instead of a build
//     variant, and its accessors always exist;
//   * a failing device call prints the message to stderr and exits with
//     EXIT_FAILURE, as checkCudaErrors does (inc/helper_cuda.h:315-334).
#ifndef ACOGPU_COLONY_H
#define ACOGPU_COLONY_H

#include <cstdint>
#include <string>

#include "acogpu.h"

// Timer slots of Colony::constructSolution (inc/Statistics.h:7-15).
enum StatsTimer {
    TIME_INIT_TOUR = 0,
    TIME_TOUR,
    TIME_COMPUTE_ANT_DISTANCE,
    TIME_EVAPORATE_PHEROMONE,
    TIME_UPDATE_PHEROMONE,
    TIME_COMPUTE_PROBABILITIES,
    TOTAL_TIMERS
};

class Colony {
   protected:
    Colony();

   public:
    // distances: n*n int32 row-major, host or device pointer, borrowed.
    Colony(int* distances, int numCities, int numAnts);
    virtual ~Colony();

    void initialize();
    void constructSolution(float timers[]);
    void computeAntDistances();
    void computeProbabilities();

    void setAlpha(float a);
    void setRho(float r);
    void setBeta(float b);
    float getAlpha();
    float getRho();
    float getBeta();
    int getNumAnts();
    int getNumCities();
    long getIterBestDist();
    long getGlobBestDist();
    int getReps();
    virtual void computeParameters() = 0;
    virtual const char* antSystemName() = 0;
    std::string getTour();  // 1-based, each city followed by ','
    void setOptimum(const int optimum) { this->optimum = optimum; }
    // Extension: up to maxIterations iterations with no host round trip per
    // iteration, stopping after the first iteration whose best reaches the
    // optimum (setOptimum) or once maxTimeMs of device time has been spent
    // (checked every 64 iterations; < 0: no limit); returns the number of
    // iterations run.
    int run(int maxIterations, float maxTimeMs = -1.0f);

    // aging (inc/Colony.h:128-139)
    int getMinRetirementAge();
    void setMinRetirementAge(int v);
    int getMaxRetirementAge();
    void setMaxRetirementAge(int v);
    int getLocalPeformance();
    void setLocalPeformance(int v);
    int getGlobalPerformance();
    void setGlobalPerformance(int v);
    int getMinPopulationLimit();
    void setMinPopulationLimit(int v);
    int getMaxPopulationLimit();
    void setMaxPopulationLimit(int v);

    unsigned int getNumValidAnts();

    // Run-time selections of this build (build variants in the reference).
    void setAging(bool on);
    bool getAging();
    void setBroker(bool on);  // accepted, no behaviour (as in the reference)
    void setCandidateListLength(int k);  // 0: every city in NN order (reference)
    void setSeed(uint64_t seed);
    aco_ctx* context() { return ctx; }

   protected:
    float greedyDistance();
    virtual void computeInitialPheromone() = 0;
    virtual void evaporatePheromones() = 0;
    virtual void updatePheromones() = 0;

    // One phase of the device iteration.  Inside constructSolution it joins
    // the iteration's sequence, which runs when the sequence is complete;
    // called on its own it runs at once.
    void runPhase(int phase);
    void check(int rc) const;

    aco_params params;
    int* distances = nullptr;
    int numCities = 0;
    int numAnts = 0;
    int reps = 0;
    int optimum = -1;
    float initialPheromone = 1e-6f;
    aco_ctx* ctx = nullptr;
    // phases named by the steps of the current constructSolution
    static const int kMaxPlan = 8;
    bool inIteration = false;
    int plan[kMaxPlan] = {};
    int planned = 0;
    // statistics of the last constructSolution (valid until the next phase)
    bool statsCached = false;
    long cachedIterBest = 0, cachedGlobBest = 0;
    unsigned int cachedValidAnts = 0;
};

#endif
