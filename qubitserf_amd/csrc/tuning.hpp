// Kernel-choice constants of the enumeration engine, measured on an MI355X.
// Regenerate with:  python -m qubitserf_amd.tune --write   (on a gfx950 device)
// Measured on AMD Instinct MI355X (gfx950), 2026-12-04: K=80 n=192 d=8 random systematic, 2.899e+10 codewords.
#pragma once

namespace qsa {
namespace tuning {

// Seconds of full-chip device time per codeword on a large level whose hot test rarely
// passes, for the VALU variants in run_tiled's table order (fd, xd, rpl): <1,1,32>
// <1,2,16> <1,2,24> <1,2,32> <2,2,16> <2,4,12> <3,4,12> <4,4,12>.
constexpr double kValuCost[8] = {8.053e-14, 8.594e-14, 7.828e-14, 8.753e-14, 1.314e-13, 1.564e-13, 2.009e-13, 2.470e-13};
// Seconds per hot-test pass (a stream step entering the medium test) of those variants.
constexpr double kValuHit[8] = {1.193e-09, 1.811e-10, 2.907e-10, 3.963e-10, 1.113e-09, 2.242e-10, 3.963e-10, 3.963e-10};
// int8 and fp4 (32x32x64) matrix-core tests: seconds per codeword, and per group of 8192
// codewords sent to the exact tests.
constexpr double kInt8Cost = 3.836e-14;
constexpr double kInt8Hit = 8.197e-10;
constexpr double kF4Cost = 4.430e-14;
constexpr double kF4Hit = 8.211e-10;
// Two chained fp4 32x32x64 products (124 test bits): per codeword, and per group of 4096.
constexpr double kF4x2Cost = 7.065e-14;
constexpr double kF4x2Hit = 8.211e-10;
// Packed fp4 32x32x64 products (30 test bits, two lane elements per row): per codeword, and
// per group of 16384.
constexpr double kF4PkCost = 2.359e-14;
constexpr double kF4PkHit = 1.420e-09;
// Fixed seconds per level launch (set-up, tail) of the VALU and matrix-core kernels.
constexpr double kValuLaunch = 3.007e-05;
constexpr double kMatrixLaunch = 2.393e-05;
// Structure factor (measured hot-test passes / fair-coin expectation): its elasticity with
// respect to the fair per-codeword pass probability (factor ~ p^-kElasticity), fitted on
// the gross code's Z side (int8 test), and the value assumed before a solve has measured it.
constexpr double kElasticity = 0.245;
constexpr double kPriorFactor = 1.0;

}  // namespace tuning
}  // namespace qsa
