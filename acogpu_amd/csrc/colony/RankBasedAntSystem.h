// RankBasedAntSystem (inc/RankBasedAntSystem.h:12-28).  The reference leaves
// evaporation and deposit empty; here they are the textbook rank-based Ant
// System (Bullnheimer, Hartl, Strauss 1999): the w-1 best ants of the
// iteration deposit (w-r)/L_r and the best-so-far tour deposits w/L_bs.
#ifndef ACOGPU_RANKBASEDANTSYSTEM_H
#define ACOGPU_RANKBASEDANTSYSTEM_H

#include "Colony.h"

class RankBasedAntSystem : public Colony {
   public:
    RankBasedAntSystem(int* distances, int cities, int ants);
    virtual ~RankBasedAntSystem();
    void initialize();
    void computeParameters();
    void setW(int w);
    int getW();
    const char* antSystemName();
    // tau0 = 0.5 w (w-1) / (rho C_nn) (src/RankBasedAntSystem.cu:44) instead
    // of the reference's constant 1e-6 when enabled.
    void setDorigoInitialPheromone(bool on);

   private:
    void computeInitialPheromone();
    void evaporatePheromones();
    void updatePheromones();
    int w;
    bool dorigoTau0 = false;
};

#endif
