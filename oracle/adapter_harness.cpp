// Test-infrastructure only: a C ABI around the drop-in adapter
// (oracle/gpu_rk2_adapter.hpp) so that tests can step a compiled-reference
// tree (handles from oracle/_ref/libbhref.so) through the GPU library.
#include <cstdio>
#include <stdexcept>

#include "gpu_rk2_adapter.hpp"
#include "ref_harness.h"

namespace {
thread_local char g_err[512];
}

extern "C" {
void* adapter_new(double cfl, const double* lo, const double* hi, int nqlt) {
    return new integ::gpu_rk2integrator(cfl, lo, hi, nqlt);
}
void adapter_free(void* a) { delete static_cast<integ::gpu_rk2integrator*>(a); }
const char* adapter_error() { return g_err; }
// 0: stepped; 1: the equation of state's domain error (state copied back);
// -1: any other error
int adapter_step(void* a, void* reftree, double* dt) {
    try {
        *dt = static_cast<integ::gpu_rk2integrator*>(a)->step(*static_cast<RefTree*>(reftree)->tree);
        return 0;
    } catch (const std::domain_error& e) {
        std::snprintf(g_err, sizeof g_err, "%s", e.what());
        return 1;
    } catch (const std::exception& e) {
        std::snprintf(g_err, sizeof g_err, "%s", e.what());
        return -1;
    }
}
}
