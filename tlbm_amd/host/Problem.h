// Problem.h -- parameters of one run, read from params.lbm.
//
// The same record and token order as the reference's Problem
// (include/Problem.h:10-45, src/Problem.cpp:5-36): 23 whitespace-separated
// tokens.  Floating-point fields are held as double (the reference holds
// them as `real`); the FP32/FP16 paths round them when they start.
#pragma once

#include <string>

class Problem {
public:
    // Reads params.lbm from the working directory.  Throws
    // std::runtime_error if it cannot be opened (as src/Problem.cpp:7-9) or
    // if a token is missing or malformed (the reference leaves such fields
    // uninitialised).
    void load_input();
    void load_input(const std::string &path);

    std::string latticeType;  // "D3Q15" | "D3Q19" | "D3Q27"
    int dynamics = 1;         // 1 = LBGK; other values: no relaxation
    int numTs = 0;
    int tsRepFreq = 1;
    int warmupTs = 0;
    int plotFreq = 1;
    double cs = 0;            // Smagorinsky constant (> 0 enables the model)
    double rhoLBM = 1;
    double uLBM = 0;
    double omega = 1;
    int nx = 0, ny = 0, nz = 0;
    int restartFlag = 0;
    int timeAvgFlag = 0;
    double lx_p = 0, ly_p = 0, lz_p = 0;
    double tConvFact = 0, lConvFact = 0, pConvFact = 0;
    int pRefIdx = 0;
    int ssDataFlag = 0;
};
