// Run loop and termination (TSP::execute src/TSP.cpp:23-171).
#include "acouav.h"

// The reference writes the summaries relative to the working directory
// (src/TSP.cpp:6, inc/def.h:80); they are skipped if the directory is absent.
static const char kSummaryFile[] = "aco/release/output/summary";

TSP::TSP(int maxTime_, int maxIter_, int maxReps_, int optimum_)
    : maxTime(maxTime_), maxIter(maxIter_), maxReps(maxReps_), optimum(optimum_) {}

// As in the reference (src/TSP.cpp:168-171) the optimum test uses the
// iteration best, and time is the summed phase time in milliseconds.
bool TSP::terminationCondition(float elapsed, int iter, int reps, long tourlen) const {
    return elapsed >= (float)maxTime || iter >= maxIter || reps >= maxReps || tourlen <= optimum;
}

void TSP::execute(const char* tspname, Colony& colony, const char* outstats, float calcdisttime, bool aging) {
    float timers[TOTAL_TIMERS];
    Writer writer(outstats ? std::string(outstats) : std::string());
    Statistics st;
    SummaryStatistics sum;
    sum.platform = "GPU";
    sum.p_alpha = colony.getAlpha();
    sum.p_beta = colony.getBeta();
    sum.p_rho = colony.getRho();
    sum.aging = aging;
    (void)tspname;

    Stopwatch init;
    init.start();
    colony.setOptimum(optimum);
    colony.initialize();
    // The reference reads the init timer without stopping it and always
    // reports 0 (src/TSP.cpp:46); the real initialisation time is reported.
    const float inittime = init.elapsed();

    int i = 0;
    for (; !terminationCondition(sum.t_totaltime, i, colony.getReps(), colony.getIterBestDist()); ++i) {
        colony.constructSolution(timers);
        st.iter = i;
        st.ants = (int)colony.getNumValidAnts();
        st.iterbest = colony.getIterBestDist();
        st.globalbest = colony.getGlobBestDist();
        st.t_inittour = timers[TIME_INIT_TOUR];
        st.t_tour = timers[TIME_TOUR];
        st.t_antdist = timers[TIME_COMPUTE_ANT_DISTANCE];
        st.t_evapphoremone = timers[TIME_EVAPORATE_PHEROMONE];
        st.t_updatephoremone = timers[TIME_UPDATE_PHEROMONE];
        st.t_computeprob = timers[TIME_COMPUTE_PROBABILITIES];
        float itertime = 0.0f;
        for (int t = 0; t < TOTAL_TIMERS; ++t) itertime += timers[t];
        sum.t_totaltime += itertime;
        sum.t_inittour += st.t_inittour;
        sum.t_computeprob += st.t_computeprob;
        sum.t_tour += st.t_tour;
        sum.t_antdist += st.t_antdist;
        sum.t_evappheromnes += st.t_evapphoremone;
        sum.t_updatephoremone += st.t_updatephoremone;
        st.t_itertime = itertime;
        st.t_globaltime = sum.t_totaltime;
        writer.write(st);
    }
    if (colony.getGlobBestDist() <= optimum) writer.write(colony.getTour());

    sum.numants = colony.getNumAnts();
    sum.numcities = colony.getNumCities();
    sum.numiter = i;
    sum.t_inittime = inittime;
    sum.t_distcalc = calcdisttime;
    sum.d_optimal = (float)optimum;
    sum.d_calulated = (float)colony.getGlobBestDist();
    sum.d_error = ((sum.d_calulated - sum.d_optimal) / sum.d_optimal) * 100;  // CALCULATE_DISTANCE_ERROR
    Writer log(kSummaryFile, ".log");
    log.writelogheader();
    log.writelog(sum);
    Writer csv(kSummaryFile, ".csv");
    csv.writecsvheader();
    csv.writecsv(sum);
}
