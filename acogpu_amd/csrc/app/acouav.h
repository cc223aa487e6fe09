// acouav: command-line front end of the MI355X solver with the reference's
// flags (src/Commandline.cpp:124-231), TSPLIB input (src/TSPReader.cu:25-100),
// per-iteration and summary output (src/Writer.cpp:81-256) and run loop
// (src/TSP.cpp:23-171).
#ifndef ACOGPU_ACOUAV_H
#define ACOGPU_ACOUAV_H

#include <chrono>
#include <cstdint>
#include <fstream>
#include <string>
#include <vector>

#include "Colony.h"

#define ACOUAV_MAX_INPUT_LEN 1024

// Wall-clock stopwatch in milliseconds (the role of Timer, src/Timer.cpp).
class Stopwatch {
   public:
    void start() { t0 = std::chrono::steady_clock::now(); }
    float elapsed() const {
        return std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

   private:
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
};

struct Statistics {  // inc/Statistics.h:18-31
    int ants = 0;
    int iter = 0;
    long iterbest = 0;
    long globalbest = 0;
    float t_inittour = 0, t_computeprob = 0, t_tour = 0, t_antdist = 0;
    float t_evapphoremone = 0, t_updatephoremone = 0, t_itertime = 0, t_globaltime = 0;
};

struct SummaryStatistics {  // inc/Statistics.h:33-53
    std::string platform;
    int numcities = 0, numants = 0, numiter = 0;
    float t_inittime = 0, t_inittour = 0, t_computeprob = 0, t_tour = 0, t_antdist = 0, t_distcalc = 0;
    float t_updatephoremone = 0, t_evappheromnes = 0, t_totaltime = 0;
    float d_optimal = 0, d_calulated = 0, d_error = 0;
    float p_alpha = 0, p_beta = 0, p_rho = 0;
    bool aging = false;
};

class Commandline {
   public:
    Commandline();
    bool parse(int argc, char* argv[]);  // false: print help
    void printHelp() const;

    std::string tsplibfile, outfile;
    int numAnts, maxTime, maxIter, maxReps, optimum;
    float alpha, beta, rho;
    bool verbose = false;
    int minRetirementAge, maxRetirementAge, localPeformance, globalPerformance;
    int minPopulationLimit, maxPopulationLimit;
    // selections that the reference makes at build time, plus extensions
    bool aging = false, broker = false, rankBased = false, dorigoTau0 = false;
    int w = 5;
    int candidates = 0;  // 0: every city in NN order, as the reference
    uint64_t seed = 12345678;
    bool hostDistances = false;  // default: device copy, as src/Setup.cpp:60-66
};

class TSPReader {
   public:
    bool read(const std::string& file);  // EUC_2D only, like the reference
    std::string name;
    int nodes = 0;
    std::vector<float> x, y;
    std::string error;
};

// EUC_2D matrix with the reference arithmetic: fp32 differences, fp64 sqrt,
// +0.5, truncation; MAX_DISTANCE on the diagonal (src/Utils.cu:162-218).
std::vector<int> euc2dMatrix(const std::vector<float>& x, const std::vector<float>& y);

class Writer {
   public:
    static bool VERBOSE;
    Writer() = default;
    explicit Writer(const std::string& file);
    Writer(const std::string& base, const char* ext);
    bool setFile(const std::string& file);
    void close();
    void writeHeader(const char* platform, float alpha, float beta, float rho, int numAnts, const std::string& aco,
                     const std::string& tsp, float calcdisttime, bool aging);
    void write(const Statistics& s);
    void write(const std::string& line);
    void writelogheader();
    void writelog(const SummaryStatistics& s);
    void writecsvheader();
    void writecsv(const SummaryStatistics& s);

   private:
    std::ofstream f;
    bool writing = false;
    std::string filename;
};

class TSP {
   public:
    TSP(int maxTime, int maxIter, int maxReps, int optimum);
    void execute(const char* tspname, Colony& colony, const char* outstats, float calcdisttime, bool aging);

   private:
    bool terminationCondition(float elapsed, int iter, int reps, long tourlen) const;
    int maxTime, maxIter, maxReps, optimum;
};

#endif
