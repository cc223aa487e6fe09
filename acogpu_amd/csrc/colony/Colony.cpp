// Colony over the C ABI.  The iteration order of constructSolution follows
// src/Colony.cu:537-594: its steps, including the virtual ones of the
// subclasses, name device phases, which run when the iteration's sequence is
// complete -- as one graph-launched iteration (aco_iterate_stats) when it is
// the standard sequence, else phase by phase -- with one wait per iteration
// for the phase times and the statistics the caller reads next.
#include "Colony.h"

#include <cstdio>
#include <cstdlib>

static const long kMaxDistance = 0x7fffffff;

Colony::Colony() { aco_params_init(&params); }

Colony::Colony(int* d, int cities, int ants) : distances(d), numCities(cities), numAnts(ants) {
    aco_params_init(&params);
    params.n = cities;
    params.m = ants;
    params.mode = ACO_MODE_FULLROW;  // the reference's candidate set: every city in NN order
    if (d == nullptr || cities <= 0 || ants <= 0) {
        std::fprintf(stderr, "Colony: invalid arguments (distances=%p, cities=%d, ants=%d)\n", (void*)d, cities,
                     ants);
        std::exit(EXIT_FAILURE);
    }
}

Colony::~Colony() {
    if (ctx) aco_destroy(ctx);
}

void Colony::check(int rc) const {
    if (rc != ACO_OK) {
        std::fprintf(stderr, "acogpu error %d: %s\n", rc, aco_last_error(ctx));
        std::exit(EXIT_FAILURE);
    }
}

void Colony::initialize() {
    computeParameters();
    params.tau0 = initialPheromone;
    if (ctx) {
        aco_destroy(ctx);
        ctx = nullptr;
    }
    statsCached = false;
    int rc = aco_create(&params, distances, nullptr, &ctx);
    if (rc == ACO_E_RANGE && params.mode == ACO_MODE_FULLROW) {
        // The reference's candidate set (every city in NN order) needs the
        // visited set and the row staged in LDS; beyond that size the colony
        // runs on candidate lists instead, and says so.
        const int k = (params.k > 0 && params.k <= 64 && params.k < params.n) ? params.k : 32;
        std::fprintf(stderr,
                     "acogpu: %s; running with candidate lists of the %d nearest cities (argmax fallback) "
                     "instead of every city in NN order\n",
                     aco_last_error(nullptr), k);
        params.mode = ACO_MODE_CANDIDATE;
        params.k = k;
        rc = aco_create(&params, distances, nullptr, &ctx);
    }
    if (rc != ACO_OK) {
        std::fprintf(stderr, "acogpu error %d: %s\n", rc, aco_last_error(nullptr));
        std::exit(EXIT_FAILURE);
    }
    check(aco_initialize(ctx));
}

void Colony::runPhase(int phase) {
    if (!ctx) {
        std::fprintf(stderr, "Colony: iteration before initialize()\n");
        std::exit(EXIT_FAILURE);
    }
    statsCached = false;
    if (inIteration) {  // part of constructSolution: runs with the whole sequence
        if (planned < kMaxPlan) plan[planned] = phase;
        ++planned;
        return;
    }
    float ms = 0.0f;
    check(aco_phase(ctx, phase, &ms));  // a step called on its own runs at once
}

void Colony::constructSolution(float timers[]) {
    // initTour (placement) and updateAntAges run inside the tour kernel and
    // the probability sweep; their slot reads 0.
    inIteration = true;
    planned = 0;
    runPhase(ACO_PHASE_PROBABILITIES);
    runPhase(ACO_PHASE_TOUR);
    runPhase(ACO_PHASE_ANT_DISTANCE);
    // virtual steps of the subclasses (AntSystem / RankBasedAntSystem)
    evaporatePheromones();
    updatePheromones();
    inIteration = false;
    static const int standard[5] = {ACO_PHASE_PROBABILITIES, ACO_PHASE_TOUR, ACO_PHASE_ANT_DISTANCE,
                                    ACO_PHASE_EVAPORATE, ACO_PHASE_UPDATE_PHEROMONE};
    bool isStandard = planned == 5;
    for (int i = 0; isStandard && i < 5; ++i) isStandard = plan[i] == standard[i];
    float t[TOTAL_TIMERS] = {0, 0, 0, 0, 0, 0};
    int64_t ib = 0, gb = 0;
    uint32_t nv = 0;
    if (isStandard) {
        check(aco_iterate_stats(ctx, t, &ib, &gb, &nv));
    } else {
        if (planned > kMaxPlan) {
            std::fprintf(stderr, "Colony: %d phases in one iteration (at most %d)\n", planned, kMaxPlan);
            std::exit(EXIT_FAILURE);
        }
        for (int i = 0; i < planned; ++i) check(aco_phase(ctx, plan[i], nullptr));
        check(aco_phase_end(ctx, t, &ib, &gb, &nv));
    }
    cachedIterBest = (long)ib;
    cachedGlobBest = (long)gb;
    cachedValidAnts = nv;
    statsCached = true;
    if (timers)
        for (int i = 0; i < TOTAL_TIMERS; ++i) timers[i] = t[i];
}

void Colony::computeAntDistances() { runPhase(ACO_PHASE_ANT_DISTANCE); }

void Colony::computeProbabilities() { runPhase(ACO_PHASE_PROBABILITIES); }

void Colony::setAlpha(float a) { params.alpha = a; }
void Colony::setRho(float r) { params.rho = r; }
void Colony::setBeta(float b) { params.beta = b; }
float Colony::getAlpha() { return params.alpha; }
float Colony::getRho() { return params.rho; }
float Colony::getBeta() { return params.beta; }
int Colony::getNumAnts() { return numAnts; }
int Colony::getNumCities() { return numCities; }
int Colony::getReps() { return reps; }

int Colony::run(int maxIterations, float maxTimeMs) {
    if (!ctx) initialize();
    statsCached = false;
    int32_t ran = 0, halted = 0;
    check(aco_run_until_time(ctx, maxIterations, optimum > 0 ? (int64_t)optimum : -1, maxTimeMs, 64, &ran, &halted,
                             nullptr));
    return ran;
}

long Colony::getIterBestDist() {
    if (!ctx) return kMaxDistance - 1;
    if (statsCached) return cachedIterBest;
    int64_t ib = 0, gb = 0;
    check(aco_get_stats(ctx, &ib, &gb, nullptr));
    return (long)ib;
}

long Colony::getGlobBestDist() {
    if (!ctx) return kMaxDistance;
    if (statsCached) return cachedGlobBest;
    int64_t ib = 0, gb = 0;
    check(aco_get_stats(ctx, &ib, &gb, nullptr));
    return (long)gb;
}

unsigned int Colony::getNumValidAnts() {
    if (!ctx) return (unsigned)numAnts;
    if (statsCached) return cachedValidAnts;
    uint32_t nv = 0;
    check(aco_get_stats(ctx, nullptr, nullptr, &nv));
    return nv;
}

std::string Colony::getTour() {
    std::string out;
    if (!ctx) return out;
    int32_t* t = new int32_t[numCities];
    check(aco_get_best_tour(ctx, t));
    for (int i = 0; i < numCities; ++i) {
        out += std::to_string(t[i] + 1);
        out += ',';
    }
    delete[] t;
    return out;
}

float Colony::greedyDistance() { return 0.0f; }  // C_nn is computed on the device by the tau0 policies

int Colony::getMinRetirementAge() { return params.min_age; }
void Colony::setMinRetirementAge(int v) { params.min_age = v; }
int Colony::getMaxRetirementAge() { return params.max_age; }
void Colony::setMaxRetirementAge(int v) { params.max_age = v; }
int Colony::getLocalPeformance() { return params.local_perf; }
void Colony::setLocalPeformance(int v) { params.local_perf = v; }
int Colony::getGlobalPerformance() { return params.global_perf; }
void Colony::setGlobalPerformance(int v) { params.global_perf = v; }
int Colony::getMinPopulationLimit() { return params.min_pop; }
void Colony::setMinPopulationLimit(int v) { params.min_pop = v; }
int Colony::getMaxPopulationLimit() { return params.max_pop; }
void Colony::setMaxPopulationLimit(int v) { params.max_pop = v; }

void Colony::setAging(bool on) { params.aging = on ? 1 : 0; }
bool Colony::getAging() { return params.aging != 0; }
void Colony::setBroker(bool on) { params.broker = on ? 1 : 0; }
void Colony::setSeed(uint64_t seed) { params.seed = seed; }

void Colony::setCandidateListLength(int k) {
    if (k <= 0) {
        params.mode = ACO_MODE_FULLROW;
    } else {
        params.mode = ACO_MODE_CANDIDATE;
        params.k = k;
    }
}
