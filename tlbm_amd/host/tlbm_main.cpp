// tlbm_main.cpp -- drop-in for the reference's driver (src/tLBM.cpp:7-88):
// the same loop, progress lines, dump schedule and LPU/s report, with the
// partition on one GPU.  Besides the reference's LPU/s (loop wall time,
// dumps included) it prints the step-only rate.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <exception>

#include "TLBM_Partition.h"

int main() {
    try {
        const int rank = 0;
        std::printf("Commencing test with %d processes\n", 1);
        TLBM_Partition myPart(0, 1);
        const int numTs = myPart.get_num_ts();
        const int tsRepFreq = myPart.get_ts_rep_freq();
        const int warmupTs = myPart.get_warmupTs();
        const int plotFreq = myPart.get_plot_freq();
        const int timeAvgFlag = myPart.get_time_avg_flag();
        // the divergence guard: every tsRepFreq steps the state is checked
        // for NaN or infinite populations (TLBM_CHECK_FINITE=0: off)
        const char *chk = std::getenv("TLBM_CHECK_FINITE");
        const bool checkFinite = !(chk && chk[0] == '0' && chk[1] == '\0');
        using clock = std::chrono::steady_clock;
        const auto timeStart = clock::now();
        double dumpSeconds = 0.0;
        for (int ts = 0; ts < numTs; ++ts) {
            const bool report = (ts + 1) % tsRepFreq == 0;
            if ((rank == 0) && report) std::printf("Executing time step: %d \n", ts + 1);
            myPart.take_LBM_time_step(ts % 2);
            if (report && checkFinite) myPart.check_finite(ts + 1);
            if (((ts + 1) % plotFreq == 0) && ((ts + 1) > warmupTs)) {
                const auto d0 = clock::now();
                myPart.write_data();
                dumpSeconds += std::chrono::duration<double>(clock::now() - d0).count();
            }
        }
        myPart.synchronize();
        const double execTime = std::chrono::duration<double>(clock::now() - timeStart).count();
        std::printf("Test complete.\n");
        std::printf("Elapsed time: %g seconds \n", execTime);
        const double numNodes = double(myPart.get_num_global_nodes());
        std::printf("Estimated LPU/s = %g \n", numNodes * double(numTs) / execTime);
        const double stepTime = execTime - dumpSeconds;
        if (stepTime > 0) std::printf("Step-only LPU/s = %g \n", numNodes * double(numTs) / stepTime);
        if (timeAvgFlag == 1) {
            std::printf("Writing time average data.\n");
            myPart.write_time_avg_data();
        }
    } catch (const std::exception &e) {
        std::fprintf(stderr, "tlbm_exec: %s\n", e.what());
        return 1;
    }
    return 0;
}
