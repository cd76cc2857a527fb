// Timing harness of the autotuner (dnsbox_amd/tune/autotune.py): the IF-Heun step of one box
// size and precision, built with a candidate configuration in place of its table entry:
//   -DDNSBOX_TUNE_N=256 -DDNSBOX_TUNE_T=float -DDNSBOX_TUNE_S=double -DDNSBOX_TUNE_CFG=CfgBase<...>
// Prints one JSON line: the in-step duration of every pass (ms), the whole-step time, and
// the energy after a few fixed-dt steps (the tuner rejects candidates whose energy differs
// from the table's beyond rounding).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>

#include "config.hpp"

using namespace dnsbox;

namespace {

dnsbox_params workload(int adaptive)
{
    dnsbox_params p;
    std::memset(&p, 0, sizeof(p));
    p.n = DNSBOX_TUNE_N;
    p.precision = sizeof(DNSBOX_TUNE_T) == 8 ? 64 : 32;
    p.state_precision = sizeof(DNSBOX_TUNE_S) == 8 ? 64 : 32;
    p.nu = 0.005 * 256.0 / double(DNSBOX_TUNE_N);
    p.kolmo_amp = 1.0;
    p.kolmo_k = 4;
    p.deltak = 1.0;
    p.cfl_min = p.cfl_max = 0.5;
    p.dt_max = 0.01;
    p.dt = 0.002;
    p.parity_mode = 1;
    p.adaptive_dt = adaptive;
    return p;
}

using Box = BoxT<Cube<DNSBOX_TUNE_N>, DNSBOX_TUNE_T, DNSBOX_TUNE_S>;

}  // namespace

int main(int argc, char** argv)
{
    const int reps = argc > 1 ? std::atoi(argv[1]) : 20;
    try {
        double ms[kNPass] = {0};
        double step_ms = 0.0;
        {
            dnsbox_params p = workload(1);
            Box box(make_config(&p));
            box.init_random(0, 0, 4.0, 0.5);
            for (int i = 0; i < 20; ++i) box.heun_step();
            box.time_passes(reps, ms);
            hipEvent_t e0, e1;
            DNSBOX_HIP(hipEventCreate(&e0));
            DNSBOX_HIP(hipEventCreate(&e1));
            DNSBOX_HIP(hipEventRecord(e0, box.stream));
            for (int i = 0; i < reps; ++i) box.heun_step();
            DNSBOX_HIP(hipEventRecord(e1, box.stream));
            DNSBOX_HIP(hipEventSynchronize(e1));
            float t = 0.f;
            DNSBOX_HIP(hipEventElapsedTime(&t, e0, e1));
            step_ms = double(t) / reps;
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        double ekin = 0.0;
        {
            dnsbox_params p = workload(0);
            Box box(make_config(&p));
            box.init_random(0, 0, 4.0, 0.5);
            for (int i = 0; i < 5; ++i) box.heun_step();
            box.spec_reduce();
            DNSBOX_HIP(hipMemcpy(&ekin, box.red_out + DG_EKIN, sizeof(double), hipMemcpyDeviceToHost));
        }
        std::printf("{\"zinv_curl\": %.6f, \"yinv6\": %.6f, \"xcross\": %.6f, \"yfwd3\": %.6f, "
                    "\"zpred_zinv\": %.6f, \"zcorr_zinv\": %.6f, \"step\": %.6f, \"ekin\": %.17g}\n",
                    ms[0], ms[1], ms[2], ms[3], ms[4], ms[5], step_ms, ekin);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "tune harness: %s\n", e.what());
        return 1;
    }
    return 0;
}
