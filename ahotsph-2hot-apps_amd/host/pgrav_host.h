// Host runtime of the batched path: octree, FP64 moments, sink groups and
// interaction lists.  These stay on the host, as in the reference (nln/mac.c,
// nln/cofm.c and libtree); the GPU only evaluates the lists.
#pragma once
#include <stdint.h>

#include <functional>
#include <string>
#include <vector>

#include "pgrav.h"

namespace pgrav {

struct TreeParams {
    double theta = 0.6;    // opening angle: accept a cell when d > bmax/theta + r_group
    int group_max = 64;    // bodies per sink group (one wave64 lane each)
    int leaf_max = 16;     // bodies per tree leaf
    int order = 4;         // highest multipole order used for cells (0, 2, 4)
    int pcut = 0;          // cells with fewer bodies than this go to the body list
                           // (0: automatic, the point where a cell costs less than its bodies)
    int periodic = 0;      // 1: nearest image (Ewald correction per pair in the kernel);
                           // 2: the (2 nimage + 1)^3 images listed explicitly (far field
                           //    by the local expansion, nln/mac.c:91-111 + nln/main.c:1811)
    int nimage = 2;        // explicit images per side for periodic = 2 (nln/main.c:246)
    double box = 1.0;      // periodic box size (origin at 0)
    int nthreads = 0;      // 0: hardware concurrency
    int merge_groups = 1;  // 0: maximal cells <= group_max; 1: adjacent small siblings merged;
                           // 2: fixed Morton runs; 3: box-limited Morton merge; 4: mode 1, then
                           // Morton-adjacent groups merged where the joint list costs less;
                           // 5: cells of <= rcb_cell bodies cut into balanced groups by
                           // coordinate bisection (groups are body sets, not runs)
    int dispatch = 2;      // 0: global longest-first; per-XCD Morton slices of equal cost with,
                           // inside, 1 longest-first, 2 Morton order, 3 costliest 2 % then Morton
    int rcb_cell = 32768;  // mode 5: cells of at most this many bodies are cut into groups
    int mixed_order = 0;   // 1: an accepted cell far enough for a lower order to meet the same
                           // truncation bound is listed at that order (order-2 / order-0 lists)
    // acceptance: 0 opening angle theta; 1 error tolerance (nln/cofm.c:276-565 critical
    // radii, nln/mac.c:1776-1800 lowest order first), the parameters of nln/main.c:533-544
    int mac = 0;
    double errtol = 0.0;      // absolute acceleration error per interaction (G = 1 units)
    double frac_tol = 0.0;    // relative to the cell's m / r^2 (0: off)
    double frac_tol0 = 0.0;   // relative floor (0: off)
    double ptol_boost = 0.0;  // errtol / (1 + ptol_boost bmax / r0)
    double stol_max = 0.0;    // > 1: errtol scaled by sigma(m) up to stol_max
    int quad_ncut = 7;        // cells of >= this many bodies have order-2 moments
    int hexa_ncut = 20;       // ... order-4 moments
    int group_box = 0;        // 1: the acceptance distance is from the group's bounding box
                              // (every sink is inside it) instead of its enclosing sphere
};

const int HOST_NXCD = 8;  // MI355X XCDs (one L2 each); workgroup b runs on XCD b % 8

struct Tree {
    int64_t n = 0;
    std::vector<int64_t> perm;    // sorted slot -> original body index
    std::vector<double> pos;      // 3n, sorted (FP64 absolute positions)
    std::vector<double> mass;     // n, sorted
    // nodes in depth-first pre-order; the children of a node are contiguous
    std::vector<int64_t> first;   // first body
    std::vector<int32_t> count;   // bodies below
    std::vector<int32_t> child;   // first child node (-1 for a leaf)
    std::vector<uint8_t> nchild;
    std::vector<uint8_t> level;
    std::vector<int32_t> parent;
    std::vector<double> cofm;     // 3 per node, FP64
    std::vector<double> bmax;     // FP64
    std::vector<double> raw;      // 35 per node: raw moments about cofm (nln/cofm.c:660-720 order)
    std::vector<double> rcrit;    // mac 1: 3 per node, critical radii of orders 0, 2, 4
                                  // (NaN where the cell has too few bodies for that order)
    double x0[3] = {0, 0, 0};
    double size = 1.0;            // root cube edge
    int64_t nnodes() const { return (int64_t)count.size(); }
};

enum ListKind { L_HEXA = 0, L_QUAD = 1, L_MONO = 2, L_BODY = 3, NLIST = 4 };

// periodic = 2: each list entry has an image, (i + n) (2n+1)^2 + (j + n) (2n+1) + (k + n)
// for the offset (i, j, k) box (nln/mac.c:91-111 numbering), in a parallel u8 array
const int MAX_NIMAGE = 2;  // (2*2+1)^3 = 125 images

struct Lists {
    int64_t ngroups = 0;
    std::vector<int64_t> gfirst;
    std::vector<int32_t> gcount;
    std::vector<double> gcenter;  // 4 per group: bounding-box centre x, y, z and enclosing radius
    std::vector<double> ghalf;    // 3 per group: bounding-box half-extents
    // mode 5: a group's bodies are gslots[gslot_off[g] .. gslot_off[g+1]) (gfirst = -1)
    std::vector<int64_t> gslot_off, gslots;
    std::vector<int64_t> off[NLIST];   // ngroups + 1 offsets per list kind
    std::vector<uint32_t> idx[NLIST];  // node indices (cells) or sorted body slots
    std::vector<uint8_t> img[NLIST];   // periodic = 2: image of each entry (else empty)
    int64_t pairs[NLIST] = {0, 0, 0, 0};  // (sink, source) pairs per kind
    std::vector<int32_t> lpt;     // dispatch order: NXCD slices, longest-first within each
    std::vector<int32_t> xoff;    // NXCD + 1 slice offsets into lpt
    std::vector<int32_t> order_np;  // one-workgroup-per-group order, slices dealt round-robin
};

// Dispatch order for a cost estimate per group (see TreeParams::dispatch).
void make_dispatch(const std::vector<double> &cost, int mode, Lists &l);

void build_tree(const double *pos, const double *mass, int64_t n, const TreeParams &p, Tree &t);
void compute_moments(Tree &t, int nthreads, const std::vector<uint64_t> *key);
void build_groups(const Tree &t, const TreeParams &p, Lists &l);
void build_lists(const Tree &t, const TreeParams &p, Lists &l);
// Prepared records for every node: FP32 (48 floats) and optionally FP64.
void prepare_cells(const Tree &t, int order, std::vector<float> &rec32, std::vector<double> *rec64,
                   std::vector<double> &cpos4);
int default_threads(int requested);
// pgrav_params -> TreeParams; an empty string, or what is wrong with the parameters
std::string tree_params_from(const pgrav_params *p, TreeParams &tp);
void parallel_for(int nthreads, int64_t n, const std::function<void(int64_t, int64_t)> &fn);
// index of the raw moment x^a y^b z^c in a node's 35
int raw_index(int a, int b, int c);
// Error-tolerance MAC (host/mac_arel.cpp): the absolute tolerance of a cell, its
// three critical radii from (m, bmax, B2, B4, B6), and every node's (Tree::rcrit).
// bmax0: the system half-extent (the reference's r0) that ptol_boost divides by.
double arel_ptol(double m, double bmax, const TreeParams &p, double bmax0);
// sb35, when given (the subtract_background form for cells of more than 8 bodies,
// nln/cofm.c:356-360, :425-434): order 2 takes B3 = sb35[0] with B4 = 0, order 4
// B5 = sb35[1] with B6 = 0.
void arel_rcrit(double m, double bmax, double B2, double B4, double B6, int64_t count,
                const TreeParams &p, double bmax0, double out[3], const double *sb35 = nullptr);
void compute_rcrit(Tree &t, const TreeParams &p, int nthreads);
// estimated VALU instructions of one wave over a group's lists (entries per kind)
double wave_cost(int64_t hexa, int64_t quad, int64_t mono, int64_t body);

// Far field of the explicit-image scheme (host/ewald_le.cpp, csrc/pgrav_le.h):
// lattice coefficient c_k, the moments-to-coefficients matrix (LE_NMON^2, row
// major) and an FP64 evaluation at the bodies (acc4 per body, G applied).
double le_lattice_coeff(int k, int nimage);
// sb: the subtract_background form (nln/ewald_le2.c:371-375, :605-646; no
// ewald_background, nln/main.c:1839): the moments less the box's uniform cube of
// the same mass, no background terms, no self term, and the acceleration (only)
// less (4 pi / 3) G / L^2 times the dipole moment about the box centre.
void le_matrix(int nimage, std::vector<double> &W, bool sb = false);
void le_evaluate_host(const double *pos3, const double *mass, int64_t n, double box, int nimage,
                      double G, double *acc4, bool sb = false);

}  // namespace pgrav
