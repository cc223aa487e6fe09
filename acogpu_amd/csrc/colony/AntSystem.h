// AntSystem (inc/AntSystem.h:7-23) over the device colony.
#ifndef ACOGPU_ANTSYSTEM_H
#define ACOGPU_ANTSYSTEM_H

#include "Colony.h"

class AntSystem : public Colony {
   public:
    AntSystem(int* distances, int numCities, int numAnts, float alpha, float beta, float rho);
    // aging variant (src/AntSystem.cu:168-181)
    AntSystem(int* distances, int numCities, int numAnts, float alpha, float beta, float rho,
              int minRetirementAge, int maxRetirementAge, int localPeformance, int globalPerformance,
              int minPopulationLimit, int maxPopulationLimit);
    virtual ~AntSystem();
    const char* antSystemName();
    void computeParameters();

   private:
    void computeInitialPheromone();
    void evaporatePheromones();
    void updatePheromones();
};

#endif
