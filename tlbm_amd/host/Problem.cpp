// Problem.cpp -- params.lbm reader (token order of src/Problem.cpp:10-32).
#include "Problem.h"

#include <fstream>
#include <stdexcept>

namespace {

template <class V>
void read_token(std::istream &in, V &v, const char *name) {
    if (!(in >> v))
        throw std::runtime_error(std::string("params.lbm: missing or malformed value for ") + name);
}

}  // namespace

void Problem::load_input() { load_input("params.lbm"); }

void Problem::load_input(const std::string &path) {
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("Could not open params.lbm!");
    read_token(in, latticeType, "latticeType");
    read_token(in, dynamics, "dynamics");
    read_token(in, numTs, "numTs");
    read_token(in, tsRepFreq, "tsRepFreq");
    read_token(in, warmupTs, "warmupTs");
    read_token(in, plotFreq, "plotFreq");
    read_token(in, cs, "cs");
    read_token(in, rhoLBM, "rhoLBM");
    read_token(in, uLBM, "uLBM");
    read_token(in, omega, "omega");
    read_token(in, nx, "nx");
    read_token(in, ny, "ny");
    read_token(in, nz, "nz");
    read_token(in, restartFlag, "restartFlag");
    read_token(in, timeAvgFlag, "timeAvgFlag");
    read_token(in, lx_p, "lx_p");
    read_token(in, ly_p, "ly_p");
    read_token(in, lz_p, "lz_p");
    read_token(in, tConvFact, "tConvFact");
    read_token(in, lConvFact, "lConvFact");
    read_token(in, pConvFact, "pConvFact");
    read_token(in, pRefIdx, "pRefIdx");
    read_token(in, ssDataFlag, "ssDataFlag");
}
