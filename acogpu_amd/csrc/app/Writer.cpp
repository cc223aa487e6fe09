// Output of acouav in the reference's formats (src/Writer.cpp:81-256): the
// header and per-iteration lines of the -p file, 12-wide left-aligned
// columns, 8 significant digits, and the summary .log/.csv rows.
#include <ctime>
#include <iomanip>
#include <iostream>

#include "acouav.h"

bool Writer::VERBOSE = false;
static const int W = 12;
static const char SEP = ',';

static bool file_empty(const std::string& name) {
    std::ifstream f(name);
    return !f.good() || f.peek() == std::ifstream::traits_type::eof();
}

Writer::Writer(const std::string& file) { setFile(file); }

Writer::Writer(const std::string& base, const char* ext) {
    if (!base.empty()) setFile(base + ext);
}

bool Writer::setFile(const std::string& file) {
    if (file.empty()) return false;
    filename = file;
    f.open(file, std::ios_base::app);
    writing = (bool)f;
    if (writing) f << std::setprecision(8);
    return writing;
}

void Writer::close() {
    if (writing) f.close();
    writing = false;
}

void Writer::writeHeader(const char* platform, float alpha, float beta, float rho, int numAnts,
                         const std::string& aco, const std::string& tsp, float calcdisttime, bool aging) {
    std::time_t now = std::time(nullptr);
    const char* date = std::ctime(&now);
    if (writing) {
        f << "\n"
          << "Date: " << date << "PLATFORM: " << platform << "\n"
          << "TSP: " << tsp << "\n"
          << "ACO: " << aco << "\n"
          << "Ants: " << numAnts << "\n"
          << "Ants Aging: " << (aging ? "TRUE" : "FALSE") << "\n"
          << "Distance calculation time: " << calcdisttime << "\n"
          << "Alpha: " << alpha << " Beta: " << beta << " Rho: " << rho << "\n"
          << std::left << std::setw(W) << "Iter#" << std::setw(W) << "#Ants" << std::setw(W) << "I-Best"
          << std::setw(W) << "G-Best" << std::setw(W) << "I-Time" << std::setw(W) << "T-Time" << std::setw(W)
          << "Init Tour" << std::setw(W) << "Tour" << std::setw(W * 2) << "Compute Ant Distances"
          << std::setw(W * 2) << "Evaporate Pheromones" << std::setw(W * 2) << "Update Pheromones"
          << std::setw(W * 2) << "Compute Probabilities" << std::endl
          << std::flush;
    }
    if (VERBOSE) {
        std::cout << "\n"
                  << "Date: " << date << "TSP: " << tsp << "\n"
                  << "ACO: " << aco << "\n"
                  << "Ants: " << numAnts << "\n"
                  << "Distance calculation time: " << calcdisttime << "\n"
                  << "Alpha: " << alpha << " Beta: " << beta << " Rho: " << rho << "\n"
                  << std::left << std::setw(W) << "Iter#" << std::setw(W) << "#Ants" << std::setw(W) << "I-Best"
                  << std::setw(W) << "G-Best" << std::setw(W) << "I-Time" << std::setw(W) << "T-Time" << std::endl;
    }
}

void Writer::write(const Statistics& s) {
    if (writing) {
        f << std::left << std::setw(W) << s.iter << std::setw(W) << s.ants << std::setw(W) << s.iterbest
          << std::setw(W) << s.globalbest << std::setw(W) << s.t_itertime << std::setw(W) << s.t_globaltime
          << std::setw(W) << s.t_inittour << std::setw(W) << s.t_tour << std::setw(W * 2) << s.t_antdist
          << std::setw(W * 2) << s.t_evapphoremone << std::setw(W * 2) << s.t_updatephoremone << std::setw(W * 2)
          << s.t_computeprob << "\n";  // buffered: the stream is flushed when the file closes
    }
    if (VERBOSE) {
        std::cout << std::left << std::setw(W) << s.iter << std::setw(W) << s.ants << std::setw(W) << s.iterbest
                  << std::setw(W) << s.globalbest << std::setw(W) << s.t_itertime << std::setw(W)
                  << s.t_globaltime << "\n";
    }
}

void Writer::write(const std::string& line) {
    if (writing) f << line << std::endl << std::flush;
}

void Writer::writelogheader() {
    if (writing && file_empty(filename)) {
        f << std::left << std::setw(W) << "platform" << std::setw(W) << "cities" << std::setw(W) << "ants"
          << std::setw(W) << "iters" << std::setw(W) << "init-t" << std::setw(W) << "caldist-t" << std::setw(W)
          << "init-tour" << std::setw(W) << "comp-prob" << std::setw(W) << "tour-time" << std::setw(W)
          << "dist-calc" << std::setw(W) << "evap-pher" << std::setw(W) << "update-pher" << std::setw(W)
          << "total-time" << std::setw(W) << "opt-dist" << std::setw(W) << "calc-dist" << std::setw(W)
          << "error %"
          << "\n"
          << std::flush;
    }
}

void Writer::writelog(const SummaryStatistics& s) {
    if (writing) {
        f << std::left << std::setw(W) << s.platform << std::setw(W) << s.numcities << std::setw(W) << s.numants
          << std::setw(W) << s.numiter << std::setw(W) << s.t_inittime << std::setw(W) << s.t_distcalc
          << std::setw(W) << s.t_inittour << std::setw(W) << s.t_computeprob << std::setw(W) << s.t_tour
          << std::setw(W) << s.t_antdist << std::setw(W) << s.t_evappheromnes << std::setw(W)
          << s.t_updatephoremone << std::setw(W) << s.t_totaltime << std::setw(W) << s.d_optimal << std::setw(W)
          << s.d_calulated << std::setw(W) << s.d_error << "\n"
          << std::flush;
    }
    if (!VERBOSE) {
        std::cout << std::left << std::setw(W) << s.platform << std::setw(W) << s.numcities << std::setw(W)
                  << s.numants << std::setw(W) << s.numiter << std::setw(W) << s.t_totaltime << std::setw(W)
                  << s.d_error << "\n";
    }
}

void Writer::writecsvheader() {
    if (writing && file_empty(filename)) {
        f << "platform" << SEP << "cities" << SEP << "ants" << SEP << "itererations" << SEP << "aging" << SEP
          << "alpha" << SEP << "beta" << SEP << "init-time" << SEP << "calc-dist-time" << SEP << "init-tour" << SEP
          << "comp-prob" << SEP << "tour-time" << SEP << "dist-calc" << SEP << "evap-pheromones" << SEP
          << "update-pheromnes" << SEP << "total-time" << SEP << "optimal-distance" << SEP << "calc-distance"
          << SEP << "error %"
          << "\n"
          << std::flush;
    }
}

void Writer::writecsv(const SummaryStatistics& s) {
    if (writing) {
        f << s.platform << SEP << s.numcities << SEP << s.numants << SEP << s.numiter << SEP
          << (s.aging ? "TRUE" : "FALSE") << SEP << s.p_alpha << SEP << s.p_beta << SEP << s.t_inittime << SEP
          << s.t_distcalc << SEP << s.t_inittour << SEP << s.t_computeprob << SEP << s.t_tour << SEP << s.t_antdist
          << SEP << s.t_evappheromnes << SEP << s.t_updatephoremone << SEP << s.t_totaltime << SEP << s.d_optimal
          << SEP << s.d_calulated << SEP << s.d_error << "\n"
          << std::flush;
    }
}
