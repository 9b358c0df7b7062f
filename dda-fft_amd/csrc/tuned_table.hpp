// Launch schedule per (lattice, right-hand sides): written by
// dda-fft_amd/tune.py from measurements on an MI355X.  x_rb: right-hand
// sides per X-pass tile; x_blocks: X-pass launch size up to which groups of
// right-hand sides are spread over blocks (1 << 30: each group its own
// block); graph: iterations replayed as captured HIP graphs; fuse: z- and
// y-forward transforms of an i-plane in one kernel.
#pragma once

namespace ddafft {

struct TunedEntry {
    int ni, nj, nl, nrhs, x_rb, x_blocks, graph, fuse;
};

// realizations in flight in ddafft_solve_batch for the batched Monte-Carlo
// workload (the call's width argument 0 stands for it)
constexpr int kTunedBatchWidth = 16;

constexpr TunedEntry kTuned[] = {
    {32, 32, 32, 2, 1, 2048, 0, 1},   // X pass 12.7 us, solve 0.913 ms
    {64, 64, 64, 2, 2, 1 << 30, 0, 1},   // X pass 49.1 us, solve 2.520 ms
    {128, 128, 128, 2, 2, 1 << 30, 0, 0},   // X pass 483.3 us, solve 9.271 ms
    {256, 256, 256, 2, 1, 1 << 30, 0, 0},   // X pass 5536.2 us, solve 94.661 ms
    {32, 32, 32, 16, 2, 8192, 0, 1},   // X pass 38.2 us, solve 2.538 ms
    {32, 32, 64, 16, 2, 1 << 30, 0, 1},   // X pass 73.0 us, solve 4.333 ms
    {64, 64, 64, 16, 2, 1 << 30, 0, 1},   // X pass 311.5 us, solve 16.310 ms
    {32, 32, 64, 2, 2, 1 << 30, 0, 1},   // X pass 18.6 us, solve 1.168 ms
    {64, 64, 64, 4, 2, 1 << 30, 0, 1},   // X pass 96.6 us, solve 4.522 ms
    {128, 128, 128, 4, 2, 1 << 30, 1, 0},   // X pass 939.5 us, solve 17.606 ms
};

}  // namespace ddafft
