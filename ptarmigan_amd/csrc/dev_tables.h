// Layout of the device-resident tables.
//
// The host loads the packed blob (ptarmigan_amd/tables/ptg_tables.bin) and
// copies the tables each kernel family needs into one contiguous "pack" with
// the compile-time offsets below. The tables read on every step come first, so
// a workgroup stages them into LDS with one coalesced loop; the event-sampling
// tables stay in global memory. Sources are the reference's embedded tables
// (file:line in ptarmigan_amd/tables/manifest.json).
#pragma once

namespace ptg {

// ---- LCFA pack (doubles)
namespace lcfa_pack {
// hot part: read every step, staged in LDS by the transport kernels
constexpr int LN_H_CHI = 0;                          // [41][2]  lcfa/photon_emission/tables.rs:9
constexpr int LN_T_CHI = LN_H_CHI + 82;              // [81][3]  lcfa/pair_creation/tables.rs:7
constexpr int GAUNT = LN_T_CHI + 243;                // [25][2]  lcfa/photon_emission/mod.rs:300
constexpr int HOT_SIZE = GAUNT + 50;                 // 375 doubles = 3 KB
// event part: read only when an emission or decay is sampled (global memory, L2-resident)
constexpr int QUANTUM_CDF = HOT_SIZE;                // [41][coeff, power, 31x2]  tables.rs:111
constexpr int Y_CDF = QUANTUM_CDF + 41 * 64;         // [31][64] tables.rs:1572
constexpr int Y_INF = Y_CDF + 31 * 64;               // [64]     tables.rs:1537
constexpr int HIGH_CHI = Y_INF + 64;                 // [32][2]  tables.rs:1467
constexpr int HIGH_CHI_PEAK = HIGH_CHI + 64;         // [32][2]  tables.rs:1502
constexpr int CLASSICAL = HIGH_CHI_PEAK + 64;        // [41][2]  tables.rs:55
constexpr int GL_NODES = CLASSICAL + 82;             // [32]     quadrature/tables.rs:111
constexpr int GL_WEIGHTS = GL_NODES + 32;            // [32]     quadrature/tables.rs:146
constexpr int AIRY_SERIES = GL_WEIGHTS + 32;         // [2][20]  special_functions/airy.rs:257
constexpr int AIRYP_SERIES = AIRY_SERIES + 40;       // [2][20]  airy.rs:423
constexpr int AIRY_N40 = AIRYP_SERIES + 40;          // nodes/weights of generalised Gauss-Laguerre rules
constexpr int AIRY_W40 = AIRY_N40 + 40;              //   (alpha = -1/6: 40, 16, 4 points;
constexpr int AIRY_N16 = AIRY_W40 + 40;              //    alpha = +1/6: 32, 16, 8, 4 points)
constexpr int AIRY_W16 = AIRY_N16 + 16;              //   airy.rs:285-421, 451-593
constexpr int AIRY_N4 = AIRY_W16 + 16;
constexpr int AIRY_W4 = AIRY_N4 + 4;
constexpr int AIRYP_N32 = AIRY_W4 + 4;
constexpr int AIRYP_W32 = AIRYP_N32 + 32;
constexpr int AIRYP_N16 = AIRYP_W32 + 32;
constexpr int AIRYP_W16 = AIRYP_N16 + 16;
constexpr int AIRYP_N8 = AIRYP_W16 + 16;
constexpr int AIRYP_W8 = AIRYP_N8 + 8;
constexpr int AIRYP_N4 = AIRYP_W8 + 8;
constexpr int AIRYP_W4 = AIRYP_N4 + 4;
constexpr int PAIR_QN = AIRYP_W4 + 4;                // [32][t, w, cosh t, cosh 2t/3, cosh(t/3)/cosh t] derived from GL_NODES
constexpr int AIRY_CHEB = PAIR_QN + 160;             // [10][t_lo, c, 1/h, 16 Chebyshev coefficients] of the
                                                     //   Gauss-Laguerre sums above (tables/gen_airy_cheb.py)
constexpr int SIZE = AIRY_CHEB + 190;                // 5953 doubles = 47.6 KB
}  // namespace lcfa_pack

// Scalars used with the LCFA pack (lcfa/photon_emission/tables.rs:4,99-103;
// lcfa/pair_creation/tables.rs:2).
struct LcfaScalars {
    double delta_ln_chi_h;     // LN_H_CHI spacing
    double ln_chi_min, ln_chi_step, ln_delta_min, ln_delta_step;
    double delta_ln_chi_t;     // LN_T_CHI spacing
};

}  // namespace ptg
