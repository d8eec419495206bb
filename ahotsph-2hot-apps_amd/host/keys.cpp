// Body keys of the reference's tree: the bounding cube (nln/physics_generic.c:77-128
// FixRsize, FixRsizeExact), the integer position (IntPos :130-144), the Morton key
// (GetKey :146-167 and its table form GetKeyFast :172-223), the Peano-Hilbert key
// (GetKeyPH :226-233 over the state machine of :340-482) and the cell corner a key
// names (CellCorner :236-258, CellCornerPH :260-272).
//
// A key is the bit string 1 x_{30} y_{30} z_{30} ... x_0 y_0 z_0 read as a binary
// number, 31 bits per axis below a leading 1 (94 bits); it is returned as two
// 64-bit words (low, high), as the reference's two-word Key_t holds it.  Float
// arithmetic follows the reference's statements (float positions, float key
// factor), so the integer positions, and with them the keys, are bit-identical.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "pgrav.h"
#include "pgrav_host.h"

namespace pgrav {
void fail(const char *fmt, ...);
void clear_error();

typedef unsigned __int128 u128;
static const int KEY_BITS_PER_AXIS = 31;  // (KEYBITS - 1) / NDIM with a 96-bit key
static const double RSIZE_EPS = 1e-7;

// The Peano-Hilbert curve as a finite automaton: a cell's state is (entry corner,
// main axis); per octant bits o = x | y << 1 | z << 2 it gives the octant's rank
// along the curve and the daughter's state.  The automaton is generated from the
// three transition sequences of nln/physics_generic.c:349-366 (which axis the
// curve crosses between consecutive octants, and which axis the daughter's own
// path runs along), as the reference's setup() does.
struct PHTables {
    int rank[4][3][8];    // octant bits -> position along the curve
    int octant[4][3][8];  // position -> octant bits
    int next_entry[4][3][8];
    int next_axis[4][3][8];
    PHTables() {
        static const int entry_mask[4] = {0, 3, 5, 6};
        static const int mask_entry[8] = {0, -1, -1, 1, -1, 2, 3, -1};
        static const int axis_mask[3] = {4, 2, 1};
        static const int cross[3][8] = {{1, 2, 1, 0, 1, 2, 1, 0},
                                        {0, 2, 0, 1, 0, 2, 0, 1},
                                        {0, 1, 0, 2, 0, 1, 0, 2}};
        static const int inner[3][8] = {{1, 2, 2, 0, 0, 2, 2, 1},
                                        {0, 2, 2, 1, 1, 2, 2, 0},
                                        {0, 1, 1, 2, 2, 1, 1, 0}};
        for (int ax = 0; ax < 3; ax++)
            for (int en = 0; en < 4; en++) {
                int corner = entry_mask[en], entry = entry_mask[en];
                for (int j = 0; j < 8; j++) {
                    rank[en][ax][corner] = j;
                    octant[en][ax][j] = corner;
                    next_entry[en][ax][corner] = mask_entry[entry];
                    next_axis[en][ax][corner] = inner[ax][j];
                    const int big = axis_mask[cross[ax][j]], small = axis_mask[inner[ax][j]];
                    corner ^= big;
                    entry ^= big ^ small;
                }
            }
    }
};
static const PHTables &ph() {
    static const PHTables t;
    return t;
}

static inline void split_key(u128 k, uint64_t *out) {
    out[0] = (uint64_t)k;
    out[1] = (uint64_t)(k >> 64);
}

static inline u128 join_key(const uint64_t *k) { return (u128)k[0] | ((u128)k[1] << 64); }

static u128 morton_key(const unsigned xp[3]) {
    u128 key = 1;
    for (int b = KEY_BITS_PER_AXIS - 1; b >= 0; b--) {
        const unsigned bits = ((xp[0] >> b) & 1) | (((xp[1] >> b) & 1) << 1) | (((xp[2] >> b) & 1) << 2);
        key = (key << 3) | bits;
    }
    return key;
}

static u128 ph_key(const unsigned xp[3]) {
    const PHTables &T = ph();
    u128 key = 1;
    int en = 0, ax = 0;
    for (int b = KEY_BITS_PER_AXIS - 1; b >= 0; b--) {
        const unsigned o = ((xp[0] >> b) & 1) | (((xp[1] >> b) & 1) << 1) | (((xp[2] >> b) & 1) << 2);
        key = (key << 3) | (u128)T.rank[en][ax][o];
        const int en2 = T.next_entry[en][ax][o];
        ax = T.next_axis[en][ax][o];
        en = en2;
    }
    return key;
}

// GetKeyFast's byte table: bit b of the byte moved to bit 3b (the table of
// nln/physics_generic.c:172-198 is this formula written out)
struct MortonByte {
    uint32_t t[256];
    MortonByte() {
        for (int i = 0; i < 256; i++) {
            uint32_t v = 0;
            for (int b = 0; b < 8; b++) v |= (uint32_t)((i >> b) & 1) << (3 * b);
            t[i] = v;
        }
    }
};

// GetKeyFast as the reference composes it, word by word, so that positions the
// fast path does not check come out as there too
static void morton_fast(const uint32_t xp[3], uint64_t *out) {
    static const MortonByte M;
    uint32_t k[4];
    for (int b = 0; b < 4; b++)
        k[b] = M.t[xp[0] >> (8 * b) & 0xff] | M.t[xp[1] >> (8 * b) & 0xff] << 1 |
               M.t[xp[2] >> (8 * b) & 0xff] << 2;
    out[0] = ((uint64_t)(k[2] & 0xffff) << 48) | ((uint64_t)k[1] << 24) | k[0];
    out[1] = (uint64_t)(1u << 29) | (uint64_t)(k[3] << 8 | k[2] >> 16);
}

// depth (levels below the root) of a key
static int key_depth(u128 key) {
    int d = 0;
    while (key > 1) {
        key >>= 3;
        d++;
    }
    return d;
}

}  // namespace pgrav

using namespace pgrav;

extern "C" {

int pgrav_host_fix_rsize(const float *rmin, const float *rmax, int exact, float *out_rmin,
                         float *rsize, float *keyfactor) {
    clear_error();
    if (!rmin || !rmax || !out_rmin || !rsize || !keyfactor) {
        fail("pgrav_host_fix_rsize: null argument");
        return -1;
    }
    float size[3], centre[3];
    for (int i = 0; i < 3; i++) {
        size[i] = rmax[i] - rmin[i];
        // FixRsize scales by the double 0.5, FixRsizeExact by 0.5f
        centre[i] = exact ? (rmax[i] + rmin[i]) * 0.5f : (float)((double)(rmax[i] + rmin[i]) * 0.5);
    }
    float d = size[0];
    if (d < size[1])
        d = size[1];
    if (d < size[2])
        d = size[2];
    const long long maxchu = 1LL << KEY_BITS_PER_AXIS;
    if (exact) {
        for (int i = 0; i < 3; i++) out_rmin[i] = -0.5f * d + centre[i];
        *keyfactor = (float)(1.0 - RSIZE_EPS) * maxchu / d;
    } else {
        d = (float)(d * (1.0 + 2. * RSIZE_EPS));
        for (int i = 0; i < 3; i++) out_rmin[i] = (float)(-0.5 * d + centre[i]);
        *keyfactor = maxchu / d;
    }
    *rsize = d;
    return 0;
}

int pgrav_host_keys(const float *pos3, int64_t n, const float *rmin, float rsize, float keyfactor,
                    int kind, uint64_t *key2) {
    clear_error();
    if ((!pos3 && n) || !rmin || (!key2 && n) || kind < 0 || kind > 2) {
        fail("pgrav_host_keys: bad arguments");
        return -1;
    }
    for (int64_t i = 0; i < n; i++) {
        const float *p = pos3 + 3 * i;
        unsigned xp[3];
        if (kind != 1) {  // IntPos: the universe check, then truncation toward zero
            for (int a = 0; a < 3; a++)
                if (p[a] < rmin[a] || p[a] > rsize + rmin[a]) {
                    fail("pgrav_host_keys: body %lld outside the key cube", (long long)i);
                    return -1;
                }
            for (int a = 0; a < 3; a++) xp[a] = (unsigned)(int)(keyfactor * (p[a] - rmin[a]));
        } else {  // GetKeyFast: straight to uint32, no check
            uint32_t xf[3];
            for (int a = 0; a < 3; a++) xf[a] = (uint32_t)(keyfactor * (p[a] - rmin[a]));
            morton_fast(xf, key2 + 2 * i);
            continue;
        }
        // GetKey and GetKeyPH read bits 30 .. 0 of the integer position
        for (int a = 0; a < 3; a++) xp[a] &= (1u << KEY_BITS_PER_AXIS) - 1;
        const u128 k = kind == 2 ? ph_key(xp) : morton_key(xp);
        split_key(k, key2 + 2 * i);
    }
    return 0;
}

int pgrav_host_cell_corner(const uint64_t *key2, int ph_curve, const float *rmin, float rsize,
                           float *corner3, float *size) {
    clear_error();
    if (!key2 || !rmin || !corner3) {
        fail("pgrav_host_cell_corner: bad arguments");
        return -1;
    }
    const u128 key = join_key(key2);
    if (key == 0) {
        fail("pgrav_host_cell_corner: key 0 names no cell");
        return -1;
    }
    const int depth = key_depth(key);
    unsigned ic[3] = {0, 0, 0};
    if (!ph_curve) {  // CellCorner: peel octants from the bottom
        u128 k = key;
        for (int d = 0; d < depth; d++, k >>= 3)
            for (int a = 0; a < 3; a++)
                if ((unsigned)(k & 7) & (1u << a))
                    ic[a] |= 1u << d;
        const float factor = rsize / (float)(1u << depth);
        for (int a = 0; a < 3; a++) corner3[a] = rmin[a] + factor * ic[a];
        if (size)
            *size = factor;
        return 0;
    }
    // CellCornerPH: walk the automaton down from the root; the corner is in units
    // of the finest cell, as the reference scales it
    const PHTables &T = ph();
    int en = 0, ax = 0;
    for (int d = depth - 1; d >= 0; d--) {
        const int j = (int)((key >> (3 * d)) & 7);
        const int o = T.octant[en][ax][j];
        for (int a = 0; a < 3; a++)
            if (o & (1 << a))
                ic[a] |= 1u << (d + KEY_BITS_PER_AXIS - depth);
        const int en2 = T.next_entry[en][ax][o];
        ax = T.next_axis[en][ax][o];
        en = en2;
    }
    const float factor = rsize / (float)(1LL << KEY_BITS_PER_AXIS);
    for (int a = 0; a < 3; a++) corner3[a] = rmin[a] + factor * ic[a];
    if (size)
        *size = factor;
    return 0;
}

}  // extern "C"
