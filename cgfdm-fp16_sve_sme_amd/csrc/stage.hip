// stage.hip — launchers of the fused RK-stage kernels.
#include "stage_general.hpp"

namespace cg {

// Every interior point, one thread each; skips the bulk box when the fast kernel covers it.
template <class S>
__global__ void __launch_bounds__(256) k_stage_general(StageParams p) {
    const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    const long long n = (long long)p.nx * p.ny * p.nz;
    if (t >= n) return;
    const int i = (int)(t % p.nx) + HALO;
    const int j = (int)((t / p.nx) % p.ny) + HALO;
    const int k = (int)(t / ((long long)p.nx * p.ny)) + HALO;
    if (p.bulk_on && i >= p.bi0 && i < p.bi1 && j >= p.bj0 && j < p.bj1 && k >= p.bk0 && k < p.bk1) return;
    stage_point_general<S>(p, i, j, k);
}

template <class S>
static int launch_stage_t(const StageParams& p, int use_fast, hipStream_t s) {
    (void)use_fast;
    StageParams q = p;
    q.bulk_on = 0;
    const long long n = (long long)p.nx * p.ny * p.nz;
    const int bs = 256;
    k_stage_general<S><<<(unsigned)((n + bs - 1) / bs), bs, 0, s>>>(q);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

int launch_stage(int prec, const StageParams& p, int use_fast, hipStream_t s) {
    switch (prec) {
        case P_F32: return launch_stage_t<float>(p, use_fast, s);
        case P_F16: return launch_stage_t<_Float16>(p, use_fast, s);
        case P_BF16: return launch_stage_t<__bf16>(p, use_fast, s);
    }
    return -1;
}

}  // namespace cg
