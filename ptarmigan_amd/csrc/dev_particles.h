// Device particle storage: structure of arrays, one array per scalar component.
#pragma once
#include <stdint.h>

namespace ptg {

// double components
enum {
    P_RT = 0, P_RX, P_RY, P_RZ,          // current position (ct, x, y, z) m
    P_UT, P_UX, P_UY, P_UZ,              // current normalised momentum
    P_R0T, P_R0X, P_R0Y, P_R0Z,          // position at creation
    P_U0T, P_U0X, P_U0Y, P_U0Z,          // momentum at creation
    P_SI, P_SQ, P_SU, P_SV,              // Stokes parameters
    P_PAYLOAD, P_PCHI, P_ICOUNT, P_WORK, P_WEIGHT,
    P_NDOUBLE
};

// meta word: bits 0-1 species, bit 2 "is the primary of its shower"
enum { META_PRIMARY = 4u };

struct PBuf {
    double* d[P_NDOUBLE];
    uint64_t* id;
    uint64_t* parent_id;
    uint64_t* stream;   // RNG stream id
    uint32_t* shower;   // index of the primary that started the shower
    uint32_t* meta;
    uint32_t* nev;      // events so far (derives the streams of secondaries)
};

constexpr size_t PBUF_BYTES_PER = P_NDOUBLE * 8 + 3 * 8 + 3 * 4;

}  // namespace ptg
