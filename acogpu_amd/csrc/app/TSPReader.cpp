// TSPLIB reader with the reference's acceptance rules (src/TSPReader.cu:25-100):
// TYPE TSP/STSP, EDGE_WEIGHT_TYPE EUC_2D, NODE_COORD_SECTION "id x y".
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>

#include "acouav.h"

static std::string strip_spaces(std::string s) {
    std::string out;
    for (char ch : s)
        if (ch != ' ' && ch != '\t') out += ch;
    return out;
}

bool TSPReader::read(const std::string& file) {
    std::ifstream in(file);
    if (!in) {
        error = "Unable to open file: " + file;
        return false;
    }
    std::string line;
    while (std::getline(in, line)) {
        if (!line.empty() && !std::isprint((unsigned char)line.back())) line.pop_back();
        if (line.size() <= 1) continue;
        std::string tag, value;
        const size_t colon = line.find(':');
        if (colon != std::string::npos) {
            tag = strip_spaces(line.substr(0, colon));
            value = strip_spaces(line.substr(colon + 1));
        } else {
            tag = strip_spaces(line);
        }
        if (tag == "NAME") {
            name = value;
        } else if (tag == "TYPE") {
            if (value != "TSP" && value != "STSP") {
                error = "Invalid problem type: " + value;
                return false;
            }
        } else if (tag == "DIMENSION") {
            nodes = std::atoi(value.c_str());
        } else if (tag == "EDGE_WEIGHT_TYPE") {
            if (value != "EUC_2D") {
                error = "Invalid edge weight type: " + value;
                return false;
            }
        } else if (tag == "NODE_COORD_SECTION") {
            x.assign(nodes, 0.0f);
            y.assign(nodes, 0.0f);
            for (int i = 0; i < nodes; ++i) {
                if (!std::getline(in, line) || line.rfind("EOF", 0) == 0) {
                    error = "NODE_COORD_SECTION ends early";
                    return false;
                }
                long id = 0;
                if (std::sscanf(line.c_str(), "%ld %f %f", &id, &x[i], &y[i]) != 3) {
                    error = "bad coordinate line: " + line;
                    return false;
                }
            }
        } else if (tag == "EOF") {
            break;
        }
    }
    if (nodes < 2 || (int)x.size() != nodes) {
        error = "no coordinates read";
        return false;
    }
    return true;
}

std::vector<int> euc2dMatrix(const std::vector<float>& x, const std::vector<float>& y) {
    const int n = (int)x.size();
    std::vector<int> d((size_t)n * n);
    for (int i = 0; i < n; ++i) {
        d[(size_t)i * n + i] = 0x7fffffff;
        for (int j = i + 1; j < n; ++j) {
            const double xd = x[i] - x[j];
            const double yd = y[i] - y[j];
            const int v = (int)(std::sqrt(xd * xd + yd * yd) + 0.5);
            d[(size_t)i * n + j] = v;
            d[(size_t)j * n + i] = v;
        }
    }
    return d;
}
