// AntSystem over the device colony (src/AntSystem.cu:159-318).
#include "AntSystem.h"

static const char kAntSystemName[] = "Ant System";

AntSystem::AntSystem(int* d, int n, int m, float alpha, float beta, float rho) : Colony(d, n, m) {
    params.variant = ACO_VARIANT_AS;
    setAlpha(alpha);
    setBeta(beta);
    setRho(rho);
}

AntSystem::AntSystem(int* d, int n, int m, float alpha, float beta, float rho, int minRetirementAge,
                     int maxRetirementAge, int localPeformance, int globalPerformance, int minPopulationLimit,
                     int maxPopulationLimit)
    : Colony(d, n, m) {
    params.variant = ACO_VARIANT_AS;
    setAlpha(alpha);
    setBeta(beta);
    setRho(rho);
    setAging(true);
    setMinRetirementAge(minRetirementAge);
    setMaxRetirementAge(maxRetirementAge);
    setLocalPeformance(localPeformance);
    setGlobalPerformance(globalPerformance);
    setMinPopulationLimit(minPopulationLimit);
    setMaxPopulationLimit(maxPopulationLimit);
}

AntSystem::~AntSystem() {}

const char* AntSystem::antSystemName() { return kAntSystemName; }

void AntSystem::computeParameters() { computeInitialPheromone(); }

// The reference uses the constant DEFAULT_PHEROMONE (src/AntSystem.cu:200-203);
// ACO_TAU0_AS (m / C_nn) is available through params.tau0_policy.
void AntSystem::computeInitialPheromone() {
    if (params.tau0_policy == ACO_TAU0_CONST) initialPheromone = params.tau0;
}

// Evaporation max((1-rho) tau, 1e-18) is applied by the probability sweep of
// the same iteration into the next pheromone buffer; this phase only checks
// the ordering and costs no kernel.
void AntSystem::evaporatePheromones() { runPhase(ACO_PHASE_EVAPORATE); }

// Symmetric deposit of 1/L by every valid ant, fp32 atomics on the device.
void AntSystem::updatePheromones() { runPhase(ACO_PHASE_UPDATE_PHEROMONE); }
