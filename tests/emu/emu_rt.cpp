// Runtime of the CPU stand-in (see hip/hip_runtime.h): workgroups in turn,
// one OS thread per lane, barriers for __syncthreads and wave exchanges.
#include "hip/hip_runtime.h"

#include <algorithm>

thread_local EmuIdx threadIdx;
EmuIdx blockIdx, gridDim, blockDim;

static std::unique_ptr<std::barrier<>> g_block_bar;
static std::vector<std::unique_ptr<std::barrier<>>> g_wave_bar;
static double g_x[16][64], g_ga[16][64], g_gb[16][64];

void emu_syncthreads() { g_block_bar->arrive_and_wait(); }

double emu_shfl_xor_d(double v, int mask) {
    const unsigned t = threadIdx.x, w = t / 64, l = t % 64;
    g_x[w][l] = v;
    g_wave_bar[w]->arrive_and_wait();
    const double r = g_x[w][l ^ (unsigned)mask];
    g_wave_bar[w]->arrive_and_wait();
    return r;
}

// lanes past the wave's end keep their own value (as on the GPU)
double emu_shfl_down_d(double v, unsigned off) {
    const unsigned t = threadIdx.x, w = t / 64, l = t % 64;
    g_x[w][l] = v;
    g_wave_bar[w]->arrive_and_wait();
    const double r = l + off < 64 ? g_x[w][l + off] : v;
    g_wave_bar[w]->arrive_and_wait();
    return r;
}

void emu_wave_gather(double a, double b, double* A, double* B) {
    const unsigned t = threadIdx.x, w = t / 64, l = t % 64;
    g_ga[w][l] = a;
    g_gb[w][l] = b;
    g_wave_bar[w]->arrive_and_wait();
    for (int i = 0; i < 64; i++) { A[i] = g_ga[w][i]; B[i] = g_gb[w][i]; }
    g_wave_bar[w]->arrive_and_wait();
}

void emu_launch(dim3 grid, dim3 block, std::function<void()> body) {
    gridDim.x = grid.x;
    blockDim.x = block.x;
    for (unsigned b = 0; b < grid.x; b++) {
        blockIdx.x = b;
        g_block_bar = std::make_unique<std::barrier<>>(block.x);
        g_wave_bar.clear();
        for (unsigned w = 0; w * 64 < block.x; w++)
            g_wave_bar.push_back(std::make_unique<std::barrier<>>(std::min(64u, block.x - 64 * w)));
        std::vector<std::thread> th;
        th.reserve(block.x);
        for (unsigned t = 0; t < block.x; t++)
            th.emplace_back([&, t] {
                threadIdx.x = t;
                body();
                g_wave_bar[t / 64]->arrive_and_drop();   // a finished lane leaves every barrier
                g_block_bar->arrive_and_drop();
            });
        for (auto& x : th) x.join();
    }
}
