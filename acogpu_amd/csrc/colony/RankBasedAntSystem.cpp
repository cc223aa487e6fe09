// RankBasedAntSystem over the device colony (src/RankBasedAntSystem.cu).
#include "RankBasedAntSystem.h"

static const char kRankName[] = "Rank-based Ant System";

RankBasedAntSystem::RankBasedAntSystem(int* d, int cities, int ants) : Colony(d, cities, ants) {
    params.variant = ACO_VARIANT_RBAS;
    w = 5;  // DEFAULT_NUM_RANK_ANTS (inc/default.h:19)
}

RankBasedAntSystem::~RankBasedAntSystem() {}

const char* RankBasedAntSystem::antSystemName() { return kRankName; }

void RankBasedAntSystem::initialize() { Colony::initialize(); }

void RankBasedAntSystem::computeParameters() {
    if (numCities < w) w = numCities;  // src/RankBasedAntSystem.cu:35-40
    params.w = w;
    computeInitialPheromone();
}

void RankBasedAntSystem::computeInitialPheromone() {
    if (dorigoTau0) {
        params.tau0_policy = ACO_TAU0_RBAS;
    } else {
        params.tau0_policy = ACO_TAU0_CONST;
        initialPheromone = params.tau0;
    }
}

void RankBasedAntSystem::setDorigoInitialPheromone(bool on) { dorigoTau0 = on; }

void RankBasedAntSystem::evaporatePheromones() { runPhase(ACO_PHASE_EVAPORATE); }

void RankBasedAntSystem::updatePheromones() { runPhase(ACO_PHASE_UPDATE_PHEROMONE); }

void RankBasedAntSystem::setW(int v) { w = v; }

int RankBasedAntSystem::getW() { return w; }
