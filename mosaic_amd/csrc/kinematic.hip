// Kinematic scattering sum on gfx950.
//
// Semantics: compute_scattering_cffi, Beam.py:663-816 -- per atom j and
// pixel p the term
//     amp_j * (f0(Q) + f' + i f'') * exp(i k (x_j + |T_p - r_j|)) * r_e
//           * [R0/r] * [sqrt(P)]
// accumulated over atoms.  The MI355X formulation:
//
//  * Staging (mosaic_kin_stage): atoms grouped by species and cut by a k-d
//    bisection into 256-atom single-species tiles.  A tile stores its origin O_t
//    in FP64 and each atom as an FP32 offset sigma_j = (r_j - O_t) / lambda
//    (in wavelengths) with its complex amplitude, 32 B per atom.
//  * Pixel prologue: per pixel, rho0 = |T_p| and Theta0 = frac(rho0/lambda)
//    in double-double, and the reference's forward-removal Q cut.
//  * Sum: per (tile, pixel) an FP64 prologue gives the origin phase
//        Theta = frac(Theta0 + (X_t + |T_p - O_t| - rho0) / lambda),
//    the direction v = R^T (x_hat - u_hat) (u_hat towards the pixel) and
//    the hoisted factor G = (f0(Q)+f'+if'') r_e [R0/rho] [sqrt P] e^{2 pi i
//    Theta}.  Per pair only FP32 work remains:
//        psi = v . sigma_j (+ second-order term when the geometry needs it),
//        S  += w_j * (cos 2 pi psi, sin 2 pi psi)
//    with sin/cos on the transcendental unit (input in revolutions), the
//    atom record read once per wave from the scalar cache (the tile address
//    is wave-uniform).  After the tile, acc += G * S in FP64.
//  * Atom slices over gridDim.y keep ~2k workgroups in flight for small
//    detectors; slice partials (FP64) are summed in slice order by a second
//    kernel, so results do not depend on scheduling.
//
// The hoist is exact to first order in (tile radius / distance); where that
// ratio exceeds 1e-6 or the dropped third-order phase term exceeds 1e-7
// revolutions, the exact per-pair kernel (k_sum_exact) runs instead.  The
// far-field forms (row recurrence, matrix-core, moment expansion) are
// described at their kernels.

#include <cstring>

#include <algorithm>
#include <type_traits>
#include <cmath>
#include <vector>

#include "common.h"
#include "pair_physics.h"

namespace mosaic {
namespace {

constexpr int TILE = 256;            // atoms per tile
constexpr int MAX_FRAMES = 65535;    // poses per call: the sum kernels' grid.z
constexpr int64_t MAX_PIXELS = int64_t(1) << 30;   // pixel indices stay 32-bit
// pixel-poses per call: the reduction and the probe launch one lane per output
constexpr int64_t MAX_OUTPUTS = (int64_t(1) << 31) - 1;
constexpr uint32_t NO_SPLIT = 0xFFFFFFFFu;   // k-d node that does not split
constexpr int SUM_THREADS = 256;     // threads per sum workgroup
constexpr int MAX_SPECIES = 16;
constexpr int TARGET_BLOCKS = 2048;  // workgroups the sum grid aims for
constexpr int MMA_P = 32;            // pixel patch edge of the matrix-core sum
constexpr int MMA_WAVES = 4;         // waves (patches) per matrix-core workgroup
constexpr int N_GEOM = 8;            // pixel geometry statistics (mosaic_kin_pixels)
constexpr int GM_MOM = 28;           // moments per (unit, patch) of the matrix-core moment sum
constexpr int GM_REC = 32;           // float2 per (unit, pose, patch) record: the moments, phase terms
constexpr int64_t GM_MAX_BYTES = int64_t(3) << 29;   // its moment buffer: at most 1.5 GiB
// The last k-d levels run inside one workgroup per node: a node of at most
// KDL_MAX atoms sits in LDS and is bisected up to KDL_LEVELS times there
// (k_kd_local<LV>, 32 << LV threads, for nodes of up to TILE << LV atoms).
constexpr int KDL_LEVELS = 5;
constexpr int KDL_MAX = TILE << KDL_LEVELS;          // 8192 atoms per node
constexpr double R_E = 2.81794092e-15;
constexpr double PI = 3.14159265358979323846;

struct AtomRec {      // 32 B, read whole by one scalar load
    float sx, sy, sz, s2;       // offset from the tile origin [wavelengths], |offset|^2
    float mag, ph;              // incident amplitude |w| and arg(w) / 2 pi
    float pad0;                 // the original atom index (bits)
    float rmag;                 // sqrt(mag), shared by the two matrix-core operands
};

struct TileRec {      // 32 B
    double ox, oy, oz;   // sample-frame origin [m]
    int species;
    int count;
};

// A global level works on chunks of KDG_CHUNK positions of one node each,
// listed by the host from the count-only plan: k_kdg_bin bins every atom's
// coordinate along its node's longest side into KDG_BINS bins and writes the
// chunk's histogram, k_kdg_select finds per node the bin holding the cut and
// each chunk's counts before it, and k_kdg_scatter writes the stable
// partition (the cut bin's first positions go left until the count is met).
// The halves overlap by at most one bin of the parent's side.
constexpr int KDG_THREADS = 256;
constexpr int KDG_PT = 8;
constexpr int KDG_CHUNK = KDG_THREADS * KDG_PT;
constexpr int KDG_BINS = 256;
static_assert(KDG_BINS == KDG_THREADS, "one bin per thread in k_kdg_select");

struct KdgNode {      // a node of a global level (heap numbering)
    uint32_t begin;   // positions [begin, begin + count)
    uint32_t count;
    uint32_t need;    // positions of the left child
    uint32_t blk0;    // its chunks [blk0, blk0 + nblk) in the level's list
    uint32_t nblk;
    uint32_t pad[3];
};

struct KdgChunk {
    uint32_t node;
    uint32_t begin;
    uint32_t end;
    uint32_t pad;
};

struct KdNode {       // a k-d node finished in LDS: positions [begin, begin + count)
    uint32_t begin;
    uint32_t count;
    uint32_t tile0;   // index of its first tile
    uint32_t species;
};

struct FrameRec {     // 128 B
    double R[9];     // sample -> lab, row-major
    double T[3];     // lab translation [m]
    double xs[3];    // R^T x_hat
    double pad;
};

// Per-pixel tables written by the pixel prologue; they sit at the head of the
// workspace and depend only on the pixel count and the number of species.
struct PixelTables {
    double* theta0;       // frac(rho0 / lambda)
    double* rho0;         // |T_p|
    double* qcut;         // forward-removal cut [1/m]
    double4* u0h;         // (T_p / rho0, 1 / (2 rho0))
    float2* gtab;         // [species][pixel] (f0+f'+if'') r_e sqrt(P), origin direction
    float4* dup;          // u0 - u0(middle pixel of the pixel's MMA_P patch)
    float2* gmma;         // [species][pixel] gtab e^{2 pi i (Theta0 - Theta0(middle))}
    double* stat_blocks;  // per prologue block: the N_GEOM statistics
    double* stats;        // reduced statistics
};

struct SpeciesConst {
    float a[5], b[5], c, f0z, fp, fpp;
};

struct PixOpts {
    double k;
    double pol_deg;
    double pol_e[3];
    int remove_forward;
    int apply_pol;
    int n_species;
    SpeciesConst spc[MAX_SPECIES];
};

struct Layout {
    PixelTables px;
    TileRec* tiles;
    AtomRec* atoms;
    float* grp_r;           // [tile][2] radius of the pair / quad headed by the tile, else +inf
    uint32_t* tile_pos;     // [tile] index in its k-d node | node's tile count << 8
    uint32_t* spc_cnt;      // species pass: [block][MAX_SPECIES] counts
    uint32_t* spc_off;      //   and each block's first slot per species
    uint32_t* vals_a;       // orders (atom index per position)
    uint32_t* vals_b;
    unsigned int* kdg_cbox; // [chunk][6] a level's tight boxes per chunk (order-preserving integers)
    float* kd_fbox[2];      // [node][6] float boxes (lo, hi) of two successive levels
    KdgNode* kdg_nodes;     // nodes of every global level
    KdgChunk* kdg_chunks;   // chunks of every global level
    uint8_t* kdg_q8;        // bin of every position
    float* kdg_xyz[2][3];   // coordinates moved with the order through the global levels
    uint32_t* kdg_bh;       // [chunk][KDG_BINS] histograms of one level
    uint2* kdg_boff;        // per chunk of one level: its node's counts below / in the cut bin before it
    uint32_t* kdg_part;     // [node][slice][KDG_BINS] slice sums of the node histograms
    int2* kdg_sel;          // per node of one level: cut bin, how many of it go left
    KdNode* kd_nodes;       // nodes handed to the in-LDS finish
    int kd_levels;
    unsigned int* radius_bits;
    float* amax_blocks;     // adjoint: per-block max |dL/dA| and max |G|
    FrameRec* frames;
    double2* partial;
    float2* gm_mom;         // moment buffer of the matrix-core moment sum
    uint8_t* gm_head;       // [tile] its unit's size if it heads one, else 0
    size_t gm_count;        // its float2 elements
    size_t total;
    int n_tiles;
    int n_slices;
    int tiles_per_slice;
    int pix_blocks;
};

constexpr int PRO_THREADS = 256;
constexpr int AMAX_BLOCKS = 256;

int pix_per_thread(int npix) { return npix >= 4 * SUM_THREADS * 64 ? 8 : 4; }

void take_pixel_tables(Carver& c, int npix, int n_species, PixelTables& P) {
    P.theta0 = c.take<double>(npix);
    P.rho0 = c.take<double>(npix);
    P.qcut = c.take<double>(npix);
    P.u0h = c.take<double4>(npix);
    P.gtab = c.take<float2>(static_cast<size_t>(n_species) * npix);
    P.dup = c.take<float4>(npix);
    P.gmma = c.take<float2>(static_cast<size_t>(n_species) * npix);
    P.stat_blocks = c.take<double>(N_GEOM * static_cast<size_t>(blocks_for(npix, PRO_THREADS)));
    P.stats = c.take<double>(N_GEOM);
}

int count_tiles(int n_species, const int64_t* counts) {
    int64_t t = 0;
    for (int s = 0; s < n_species; ++s) t += (counts[s] + TILE - 1) / TILE;
    return static_cast<int>(t);
}

// Workspace layout; also the size query when base is null.
// Levels of the k-d bisection: enough to bring the largest species down to
// one tile per leaf.
int kd_levels(int n_species, const int64_t* counts) {
    int64_t tmax = 1;
    for (int s = 0; s < n_species; ++s) tmax = std::max<int64_t>(tmax, (counts[s] + TILE - 1) / TILE);
    int levels = 0;
    while ((int64_t(1) << levels) < tmax) ++levels;
    return levels;
}

// Node ranges of the k-d bisection for levels 0..levels, in heap numbering
// (node h of a level has children 2h and 2h + 1 on the next): a node of n > T
// atoms keeps ceil(n / 2T) T on the left; a node of one tile or less passes
// whole to its left child.
void kd_plan_levels(int n_species, const int64_t* counts, int levels,
                    std::vector<std::vector<int64_t>>& B, std::vector<std::vector<int64_t>>& E) {
    B.assign(levels + 1, {});
    E.assign(levels + 1, {});
    int64_t a = 0;
    for (int s = 0; s < n_species; ++s) {
        B[0].push_back(a);
        a += counts[s];
        E[0].push_back(a);
    }
    for (int l = 0; l < levels; ++l) {
        for (size_t h = 0; h < B[l].size(); ++h) {
            const int64_t b = B[l][h], e = E[l][h], n = e - b;
            const int64_t cut = n > TILE ? b + (n + 2 * TILE - 1) / (2 * TILE) * TILE : e;
            B[l + 1].push_back(b);
            E[l + 1].push_back(cut);
            B[l + 1].push_back(cut);
            E[l + 1].push_back(e);
        }
    }
}

// Global levels: the k-d levels above the last KDL_LEVELS.
int kd_global_levels(int n_species, const int64_t* counts) {
    return std::max(0, kd_levels(n_species, counts) - KDL_LEVELS);
}

// Slices over which k_kdg_total sums a level's node histograms: about 16
// chunks each for the level's largest node.
int kdg_slices(const std::vector<KdgNode>& nodes, size_t n0, size_t n1) {
    uint32_t mx = 1;
    for (size_t i = n0; i < n1; ++i) mx = std::max(mx, nodes[i].nblk);
    return static_cast<int>(std::min<uint32_t>(64u, (mx + 15u) / 16u));
}

// Workspace sizes of the global levels' tables from the plan alone: nodes
// and chunks over all levels, and per level the largest chunk count, node
// count and node x slice count (kdg_slices).
struct KdgSizes {
    size_t nodes = 0, chunks = 0, max_chunks = 1, max_nodes = 1, max_part = 1;
};
KdgSizes kdg_sizes(int n_species, const int64_t* counts) {
    const int G = kd_global_levels(n_species, counts);
    std::vector<std::vector<int64_t>> B, E;
    kd_plan_levels(n_species, counts, G, B, E);
    KdgSizes z;
    for (int l = 0; l < G; ++l) {
        size_t ch = 0;
        uint32_t mx = 1;
        for (size_t h = 0; h < B[l].size(); ++h) {
            const uint32_t nb = static_cast<uint32_t>((E[l][h] - B[l][h] + KDG_CHUNK - 1) / KDG_CHUNK);
            ch += nb;
            mx = std::max(mx, nb);
        }
        const size_t S = std::min<uint32_t>(64u, (mx + 15u) / 16u);
        z.nodes += B[l].size();
        z.chunks += ch;
        z.max_chunks = std::max(z.max_chunks, ch);
        z.max_nodes = std::max(z.max_nodes, B[l].size());
        z.max_part = std::max(z.max_part, B[l].size() * S);
    }
    return z;
}

// Node and chunk tables of the global levels; per level the offsets of its
// nodes and chunks in them.
void kdg_tables(int n_species, const int64_t* counts, std::vector<KdgNode>& nodes,
                std::vector<KdgChunk>& chunks, std::vector<size_t>& node_off,
                std::vector<size_t>& chunk_off) {
    const int G = kd_global_levels(n_species, counts);
    std::vector<std::vector<int64_t>> B, E;
    kd_plan_levels(n_species, counts, G, B, E);
    nodes.clear();
    chunks.clear();
    node_off.assign(G + 1, 0);
    chunk_off.assign(G + 1, 0);
    for (int l = 0; l < G; ++l) {
        node_off[l] = nodes.size();
        chunk_off[l] = chunks.size();
        for (size_t h = 0; h < B[l].size(); ++h) {
            const int64_t b = B[l][h], n = E[l][h] - b;
            KdgNode nd = {};
            nd.begin = static_cast<uint32_t>(b);
            nd.count = static_cast<uint32_t>(n);
            nd.need = static_cast<uint32_t>(n > TILE ? (n + 2 * TILE - 1) / (2 * TILE) * TILE : n);
            nd.blk0 = static_cast<uint32_t>(chunks.size() - chunk_off[l]);
            for (int64_t c0 = b; c0 < b + n; c0 += KDG_CHUNK) {
                KdgChunk c = {static_cast<uint32_t>(h), static_cast<uint32_t>(c0),
                              static_cast<uint32_t>(std::min<int64_t>(c0 + KDG_CHUNK, b + n)), 0u};
                chunks.push_back(c);
            }
            nd.nblk = static_cast<uint32_t>(chunks.size() - chunk_off[l]) - nd.blk0;
            nodes.push_back(nd);
        }
    }
    node_off[G] = nodes.size();
    chunk_off[G] = chunks.size();
}

int make_layout(void* base, int64_t n_atoms, int n_species, const int64_t* counts,
                int n_frames, int Ny, int Nz, Layout& L) {
    const int npix = Ny * Nz;
    Carver c(base);
    take_pixel_tables(c, npix, n_species, L.px);
    L.n_tiles = count_tiles(n_species, counts);
    L.tiles = c.take<TileRec>(std::max(1, L.n_tiles));
    L.atoms = c.take<AtomRec>(static_cast<size_t>(std::max(1, L.n_tiles)) * TILE);
    L.grp_r = c.take<float>(2 * static_cast<size_t>(std::max(1, L.n_tiles)));
    L.tile_pos = c.take<uint32_t>(static_cast<size_t>(std::max(1, L.n_tiles)));
    size_t n = static_cast<size_t>(std::max<int64_t>(1, n_atoms));
    L.vals_a = c.take<uint32_t>(n);
    L.vals_b = c.take<uint32_t>(n);
    {
        const size_t nblk = (n + KDG_CHUNK - 1) / KDG_CHUNK;
        L.spc_cnt = c.take<uint32_t>(nblk * MAX_SPECIES);
        L.spc_off = c.take<uint32_t>(nblk * MAX_SPECIES);
    }
    L.kd_levels = kd_levels(n_species, counts);
    const size_t top = static_cast<size_t>(n_species) << std::max(0, L.kd_levels - 1);
    L.kd_fbox[0] = c.take<float>(12 * top);
    L.kd_fbox[1] = c.take<float>(12 * top);
    {
        const KdgSizes z = kdg_sizes(n_species, counts);
        // the staging plan's three tables side by side (one upload)
        L.kdg_nodes = c.take<KdgNode>(std::max<size_t>(1, z.nodes));
        L.kdg_chunks = c.take<KdgChunk>(std::max<size_t>(1, z.chunks));
        L.kd_nodes = c.take<KdNode>(std::max(1, L.n_tiles));
        L.kdg_q8 = c.take<uint8_t>(n);
        for (int b = 0; b < 2; ++b)
            for (int d = 0; d < 3; ++d) L.kdg_xyz[b][d] = c.take<float>(n);
        L.kdg_bh = c.take<uint32_t>(z.max_chunks * KDG_BINS);
        L.kdg_boff = c.take<uint2>(z.max_chunks);
        L.kdg_cbox = c.take<unsigned int>(6 * z.max_chunks);
        L.kdg_part = c.take<uint32_t>(z.max_part * KDG_BINS);
        L.kdg_sel = c.take<int2>(z.max_nodes);
    }
    L.radius_bits = c.take<unsigned int>(2);
    L.amax_blocks = c.take<float>(2 * AMAX_BLOCKS);
    L.frames = c.take<FrameRec>(std::max(1, n_frames));
    const int ppt = pix_per_thread(npix);
    L.pix_blocks = blocks_for(npix, SUM_THREADS * ppt);
    // the smaller of the pixel grids of the row kernels and of the
    // matrix-core sum (one wave per 32 x 32 patch, MMA_WAVES per workgroup)
    const int mma_blocks = std::max(1, npix / (MMA_P * MMA_P * MMA_WAVES));
    const int64_t grid_px =
        static_cast<int64_t>(std::min(L.pix_blocks, mma_blocks)) * std::max(1, n_frames);
    int slices = static_cast<int>((TARGET_BLOCKS + grid_px - 1) / grid_px);
    // Each slice writes FP64 partials for every pixel and pose (16 B) that the
    // reduction reads back: keep them within four times the atom records the
    // slice reads (32 B per atom), at least npix * poses / 2048 tiles per
    // slice, so that a small detector does not spend more on partials than on
    // the sum (a 128^2 frame of 200k atoms: 512 slices of 1.5 tiles before).
    const int64_t min_tps = static_cast<int64_t>(npix) * std::max(1, n_frames) / 2048;
    if (min_tps > 1)
        slices = static_cast<int>(std::min<int64_t>(
            slices, (std::max(1, L.n_tiles) + min_tps - 1) / min_tps));
    slices = std::max(1, std::min(slices, std::max(1, L.n_tiles)));
    L.tiles_per_slice = (std::max(1, L.n_tiles) + slices - 1) / slices;
    L.n_slices = (std::max(1, L.n_tiles) + L.tiles_per_slice - 1) / L.tiles_per_slice;
    L.partial = c.take<double2>(static_cast<size_t>(L.n_slices) * std::max(1, n_frames) * npix);
    // the matrix-core moment sum's buffer: a GM_REC record per (unit head, pose,
    // 32 x 32 patch), for every tile if it fits, else for rounds of tiles;
    // and the unit size of every head tile (0 elsewhere)
    {
        const int64_t patches = static_cast<int64_t>((Ny + MMA_P - 1) / MMA_P) * ((Nz + MMA_P - 1) / MMA_P);
        const int64_t per_tile = patches * std::max(1, n_frames) * GM_REC;
        const int64_t cap = std::max<int64_t>(per_tile, GM_MAX_BYTES / static_cast<int64_t>(sizeof(float2)));
        L.gm_count = static_cast<size_t>(std::min<int64_t>(per_tile * std::max(1, L.n_tiles), cap));
        L.gm_mom = c.take<float2>(L.gm_count);
        L.gm_head = c.take<uint8_t>(static_cast<size_t>(std::max(1, L.n_tiles)));
    }
    L.total = c.off + 256;
    return MOSAIC_OK;
}

// ----------------------------------------------------------------- reductions
__device__ __forceinline__ double wave_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Block-wide min of one value per thread (blockDim multiple of 64, <= 1024).
__device__ double block_min(double v, double* lds) {
    v = wave_min(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) lds[w] = v;
    __syncthreads();
    const int nw = blockDim.x >> 6;
    double r = lds[0];
    for (int i = 1; i < nw; ++i) r = fmin(r, lds[i]);
    __syncthreads();
    return r;
}
__device__ double block_max(double v, double* lds) {
    v = wave_max(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) lds[w] = v;
    __syncthreads();
    const int nw = blockDim.x >> 6;
    double r = lds[0];
    for (int i = 1; i < nw; ++i) r = fmax(r, lds[i]);
    __syncthreads();
    return r;
}

// ------------------------------------------------------------- pixel prologue
// Unit vector from the lab origin to pixel p.
__device__ __forceinline__ void pixel_dir(const double* cx, const double* cy, const double* cz,
                                          int p, double& ux, double& uy, double& uz) {
    const double tx = cx[p], ty = cy[p], tz = cz[p];
    const double r = sqrt(tx * tx + ty * ty + tz * tz);
    const double inv = r > 0.0 ? 1.0 / r : 0.0;
    ux = tx * inv;
    uy = ty * inv;
    uz = tz * inv;
}

// Slope g and curvature h of the quadratic f(a) - f(0) = g a + h a^2 through
// the values d1 = f(a1) - f(0) and d2 = f(a2) - f(0), a1 <= 0 <= a2; a line
// when a2 = 0 (two points), zero for a single point.
__device__ __forceinline__ void quad_fit(double d1, double d2, int a1, int a2, double& g,
                                         double& h) {
    if (a1 == 0) { g = 0.0; h = 0.0; return; }
    if (a2 == 0) { g = d1 / a1; h = 0.0; return; }
    h = (d2 / a2 - d1 / a1) / (a2 - a1);
    g = d1 / a1 - h * a1;
}

// The matrix-core sum's model of a direction field over one MMA_P x MMA_P
// pixel patch: f(a, b) = f_c + a gy + a^2 hy + b gz + b^2 hz, a and b the
// offsets from the patch's middle pixel (my, mz) = (vy / 2, vz / 2), fitted
// through the middle pixel and the ends of the middle row and column.
struct PatchGeom {
    int y0, z0, vy, vz, my, mz;
};
__device__ __forceinline__ PatchGeom patch_of(int py, int pz, int Ny, int Nz) {
    PatchGeom g;
    g.y0 = py * MMA_P;
    g.z0 = pz * MMA_P;
    g.vy = min(MMA_P, Ny - g.y0);
    g.vz = min(MMA_P, Nz - g.z0);
    g.my = g.vy / 2;
    g.mz = g.vz / 2;
    return g;
}

// Per pixel:
//  * rho0 = |T| and Theta0 = frac(K rho0) in double-double (K = k / 2 pi);
//  * the forward-removal cut-off from the right/upper neighbours
//    (Beam.py:718-746);
//  * u0 = T / rho0 and 1 / (2 rho0), for the far-field expansion of the
//    (tile, pixel) origin phase;
//  * per species, G = (f0(Q) + f' + i f'' [- forward term]) r_e sqrt(P) in
//    the direction u0 (Beam.py:748-807 with the atom at the lab origin);
//  * block statistics: nearest pixel distance, the largest first and second
//    differences of u0 between neighbours along a row and a column, the
//    largest deviation of u0 from the separable patch model of the
//    matrix-core sum and from the moment sum's model with the cross term
//    a b g_yz, and the largest |g_yz|.
__global__ __launch_bounds__(PRO_THREADS) void k_pixels(
    const double* __restrict__ cx, const double* __restrict__ cy,
    const double* __restrict__ cz, int Ny, int Nz, const PixOpts O, PixelTables T) {
    __shared__ double lds[PRO_THREADS / 64];
    const int npix = Ny * Nz;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const double k = O.k;
    double rmin = 1e300, du1 = 0.0, du2 = 0.0, dpatch = 0.0, du2z = 0.0, du1z = 0.0;
    double dpatchx = 0.0, gyzn = 0.0;
    if (p < npix) {
        const double tx = cx[p], ty = cy[p], tz = cz[p];
        double f, r, r_lo;
        origin_phase(tx, ty, tz, k, f, r, r_lo);
        T.theta0[p] = f;
        T.rho0[p] = r + r_lo;
        rmin = r;

        const int ix = p % Ny, iy = p / Ny;
        double ux = 0.0, uy = 0.0, uz = 0.0;
        if (r > 0.0) { ux = tx / r; uy = ty / r; uz = tz / r; }
        const int nr = (ix + 1 < Ny) ? (p + 1) : ((ix > 0) ? (p - 1) : p);
        const int nu = (iy + 1 < Nz) ? (p + Ny) : ((iy > 0) ? (p - Ny) : p);
        double q2[2];
        const int nb[2] = {nr, nu};
        for (int t = 0; t < 2; ++t) {
            const double rx = cx[nb[t]], ry = cy[nb[t]], rz = cz[nb[t]];
            const double Rr = sqrt(rx * rx + ry * ry + rz * rz);
            double cd = (Rr > 0.0) ? (ux * rx + uy * ry + uz * rz) / Rr : 1.0;
            cd = fmin(1.0, fmax(-1.0, cd));
            q2[t] = k * sqrt(fmax(0.0, 2.0 * (1.0 - cd)));
        }
        const double qcut = 0.5 * sqrt(q2[0] * q2[0] + q2[1] * q2[1]);
        T.qcut[p] = qcut;
        T.u0h[p] = make_double4(ux, uy, uz, r > 0.0 ? 0.5 / r : 0.0);

        // form factor, polarization in the origin direction
        const double omc = (tx > 0.0) ? (ty * ty + tz * tz) / (r * (r + tx)) : 1.0 - ux;
        const double Q = k * sqrt(fmax(0.0, 2.0 * omc));
        const double sq = Q * (0.25e-10 / PI);
        const double ss = sq * sq;
        double scale = R_E;
        if (O.apply_pol) {
            const double er = O.pol_e[0] * ux + O.pol_e[1] * uy + O.pol_e[2] * uz;
            double P = 0.5 * (1.0 - O.pol_deg) * (1.0 + ux * ux) + O.pol_deg * (1.0 - er * er);
            P = fmin(1.0, fmax(0.0, P));
            scale *= sqrt(P);
        }
        // the pixel's matrix-core patch and its middle pixel
        const PatchGeom G = patch_of(ix / MMA_P, iy / MMA_P, Ny, Nz);
        const int pm = (G.z0 + G.mz) * Ny + G.y0 + G.my;
        double mc[3];
        pixel_dir(cx, cy, cz, pm, mc[0], mc[1], mc[2]);
        T.dup[p] = make_float4(static_cast<float>(ux - mc[0]), static_cast<float>(uy - mc[1]),
                               static_cast<float>(uz - mc[2]), 0.f);
        double fm, rm, rm_lo;
        origin_phase(cx[pm], cy[pm], cz[pm], k, fm, rm, rm_lo);
        double dsn, dcs;
        sincospi(2.0 * (f - fm), &dsn, &dcs);

        const bool cut = O.remove_forward && Q < qcut;
        for (int sp = 0; sp < O.n_species; ++sp) {
            const SpeciesConst& S = O.spc[sp];
            double f0 = S.c;
            for (int m = 0; m < 5; ++m) f0 += static_cast<double>(S.a[m]) * exp(-static_cast<double>(S.b[m]) * ss);
            double sr = f0 + S.fp, si = S.fpp;
            if (cut) { sr -= static_cast<double>(S.f0z) + S.fp; si -= S.fpp; }
            T.gtab[static_cast<int64_t>(sp) * npix + p] =
                make_float2(static_cast<float>(sr * scale), static_cast<float>(si * scale));
            T.gmma[static_cast<int64_t>(sp) * npix + p] =
                make_float2(static_cast<float>((sr * dcs - si * dsn) * scale),
                            static_cast<float>((sr * dsn + si * dcs) * scale));
        }

        // direction differences along the row
        if (ix + 1 < Ny) {
            double vx, vy, vz;
            pixel_dir(cx, cy, cz, p + 1, vx, vy, vz);
            du1 = sqrt((vx - ux) * (vx - ux) + (vy - uy) * (vy - uy) + (vz - uz) * (vz - uz));
            if (ix > 0) {
                double wx, wy, wz;
                pixel_dir(cx, cy, cz, p - 1, wx, wy, wz);
                const double ax = vx - 2.0 * ux + wx, ay = vy - 2.0 * uy + wy,
                             az = vz - 2.0 * uz + wz;
                du2 = sqrt(ax * ax + ay * ay + az * az);
            }
        }
        // and along the column
        if (iy + 1 < Nz) {
            double vx, vy, vz;
            pixel_dir(cx, cy, cz, p + Ny, vx, vy, vz);
            du1z = sqrt((vx - ux) * (vx - ux) + (vy - uy) * (vy - uy) + (vz - uz) * (vz - uz));
            if (iy > 0) {
                double wx, wy, wz;
                pixel_dir(cx, cy, cz, p - Ny, wx, wy, wz);
                const double ax = vx - 2.0 * ux + wx, ay = vy - 2.0 * uy + wy,
                             az = vz - 2.0 * uz + wz;
                du2z = sqrt(ax * ax + ay * ay + az * az);
            }
        }

        // deviation from the patch model
        double e1[3], e2[3], e3[3], e4[3];
        pixel_dir(cx, cy, cz, pm - G.my, e1[0], e1[1], e1[2]);
        pixel_dir(cx, cy, cz, pm - G.my + G.vy - 1, e2[0], e2[1], e2[2]);
        pixel_dir(cx, cy, cz, pm - G.mz * Ny, e3[0], e3[1], e3[2]);
        pixel_dir(cx, cy, cz, pm + (G.vz - 1 - G.mz) * Ny, e4[0], e4[1], e4[2]);
        const int oa = ix - G.y0 - G.my, ob = iy - G.z0 - G.mz;
        const double u0v[3] = {ux, uy, uz};
        // and from the model with the cross term a b g_yz of the moment sum,
        // g_yz the mixed difference of the patch's corners
        double c00[3], c01[3], c10[3], c11[3];
        const int pc = pm - G.my - G.mz * Ny;
        pixel_dir(cx, cy, cz, pc, c00[0], c00[1], c00[2]);
        pixel_dir(cx, cy, cz, pc + G.vy - 1, c01[0], c01[1], c01[2]);
        pixel_dir(cx, cy, cz, pc + (G.vz - 1) * Ny, c10[0], c10[1], c10[2]);
        pixel_dir(cx, cy, cz, pc + G.vy - 1 + (G.vz - 1) * Ny, c11[0], c11[1], c11[2]);
        const double area = static_cast<double>(G.vy - 1) * static_cast<double>(G.vz - 1);
        double dev2 = 0.0, devx2 = 0.0, g2 = 0.0;
        for (int d = 0; d < 3; ++d) {
            double gy, hy, gz, hz;
            quad_fit(e1[d] - mc[d], e2[d] - mc[d], -G.my, G.vy - 1 - G.my, gy, hy);
            quad_fit(e3[d] - mc[d], e4[d] - mc[d], -G.mz, G.vz - 1 - G.mz, gz, hz);
            const double model = mc[d] + oa * (gy + oa * hy) + ob * (gz + ob * hz);
            dev2 += (u0v[d] - model) * (u0v[d] - model);
            const double gyz = area > 0.0 ? (c11[d] - c10[d] - c01[d] + c00[d]) / area : 0.0;
            const double mx = model + static_cast<double>(oa) * ob * gyz;
            devx2 += (u0v[d] - mx) * (u0v[d] - mx);
            g2 += gyz * gyz;
        }
        dpatch = sqrt(dev2);
        dpatchx = sqrt(devx2);
        gyzn = sqrt(g2);
    }
    rmin = block_min(rmin, lds);
    du1 = block_max(du1, lds);
    du2 = block_max(du2, lds);
    dpatch = block_max(dpatch, lds);
    du2z = block_max(du2z, lds);
    du1z = block_max(du1z, lds);
    dpatchx = block_max(dpatchx, lds);
    gyzn = block_max(gyzn, lds);
    if (threadIdx.x == 0) {
        T.stat_blocks[N_GEOM * blockIdx.x] = rmin;
        T.stat_blocks[N_GEOM * blockIdx.x + 1] = du1;
        T.stat_blocks[N_GEOM * blockIdx.x + 2] = du2;
        T.stat_blocks[N_GEOM * blockIdx.x + 3] = dpatch;
        T.stat_blocks[N_GEOM * blockIdx.x + 4] = du2z;
        T.stat_blocks[N_GEOM * blockIdx.x + 5] = du1z;
        T.stat_blocks[N_GEOM * blockIdx.x + 6] = dpatchx;
        T.stat_blocks[N_GEOM * blockIdx.x + 7] = gyzn;
    }
}

__global__ void k_stats_reduce(const double* __restrict__ in, int n, double* __restrict__ out) {
    __shared__ double lds[16];
    double v[N_GEOM] = {1e300, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        v[0] = fmin(v[0], in[N_GEOM * i]);
        for (int q = 1; q < N_GEOM; ++q) v[q] = fmax(v[q], in[N_GEOM * i + q]);
    }
    v[0] = block_min(v[0], lds);
    for (int q = 1; q < N_GEOM; ++q) v[q] = block_max(v[q], lds);
    if (threadIdx.x == 0)
        for (int q = 0; q < N_GEOM; ++q) out[q] = v[q];
}

// ------------------------------------------------------------------- staging
// Tiles come from a k-d bisection of each species' atoms.  A node of n atoms
// is cut across the longest side of its bounding box so that the left part
// holds ceil(n / 2T) T atoms (T = 256); after ceil(log2(tiles)) levels every
// leaf is one tile, all full except the last of each species, in the same
// places as a plain cut of the sorted order into runs of T.  The cells are
// boxes of near-equal count and bounded aspect, so the tile radius stays near
// that of a cube of 256 atoms whatever the crystal's shape or orientation (a
// space-filling-curve order leaves runs that wander through the empty corners
// of the bounding box: radius up to ~120 wavelengths instead of ~20).  The
// global levels are one key pass (node, coordinate along the node's longest
// side) and one stable radix sort over all nodes at once; the last KDL_LEVELS
// levels and the tile records run in LDS, one workgroup per node
// (k_kd_local).

// Global levels whose boxes are measured over the atoms: the first and the
// last KD_TIGHT_LEVELS; the others cut their parents' boxes at the cut bin.
// Loose boxes near the leaves let an elongated node through (on the 5M-atom
// perovskite the largest tile grows from 22 to 28 wavelengths), tight ones
// there keep it near 22.
constexpr int KD_TIGHT_LEVELS = 2;


__device__ __forceinline__ unsigned int f2ord(float f) {
    const unsigned int b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float ord2f(unsigned int u) {
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}
__device__ __forceinline__ unsigned int wave_umin(unsigned int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, static_cast<unsigned int>(__shfl_xor(static_cast<int>(v), o, 64)));
    return v;
}
__device__ __forceinline__ unsigned int wave_umax(unsigned int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, static_cast<unsigned int>(__shfl_xor(static_cast<int>(v), o, 64)));
    return v;
}

__global__ void k_stage_init(unsigned int* __restrict__ radius_bits) {
    if (threadIdx.x < 2) radius_bits[threadIdx.x] = 0u;
}

constexpr int KD_THREADS = 256;

// Longest side of a float box (lo[3], hi[3]); ties go to the lower axis.
__device__ __forceinline__ int kd_axis(const float* b, float (&lo)[3], float (&ext)[3]) {
    for (int d = 0; d < 3; ++d) {
        lo[d] = b[d];
        ext[d] = b[3 + d] - lo[d];
    }
    int ax = 0;
    if (ext[1] > ext[ax]) ax = 1;
    if (ext[2] > ext[ax]) ax = 2;
    return ax;
}

// The global levels move each position's coordinates and atom index
// together, in four arrays (k_species_scatter fills the first set), so that
// every level reads and writes them contiguously.
struct KdgSet {
    float* x;
    float* y;
    float* z;
    uint32_t* idx;
};

// Global level: tight boxes, one chunk per workgroup, each chunk's own box
// (order-preserving images of the floats; lo = ~0, hi = 0 for no positions)
// into cbox[chunk]; k_kdg_box_node combines a node's chunks.  (One atomic
// min / max per chunk on the node's words had every chunk of a level with
// few nodes contend for the same six words.)
__global__ __launch_bounds__(KDG_THREADS) void k_kdg_box(
    const KdgSet in, const KdgChunk* __restrict__ chunks, unsigned int* __restrict__ cbox) {
    __shared__ unsigned int red[KDG_THREADS / 64][6];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const KdgChunk C = chunks[blockIdx.x];
    unsigned int lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
#pragma unroll
    for (int r = 0; r < KDG_PT; ++r) {
        const uint32_t i = C.begin + tid + r * KDG_THREADS;
        if (i < C.end) {
            const unsigned int o[3] = {f2ord(in.x[i]), f2ord(in.y[i]), f2ord(in.z[i])};
            for (int d = 0; d < 3; ++d) {
                lo[d] = min(lo[d], o[d]);
                hi[d] = max(hi[d], o[d]);
            }
        }
    }
    for (int d = 0; d < 3; ++d) {
        lo[d] = wave_umin(lo[d]);
        hi[d] = wave_umax(hi[d]);
    }
    if (lane == 0)
        for (int d = 0; d < 3; ++d) { red[wave][d] = lo[d]; red[wave][3 + d] = hi[d]; }
    __syncthreads();
    if (tid < 6) {
        unsigned int v = red[0][tid];
        for (int w = 1; w < KDG_THREADS / 64; ++w)
            v = tid < 3 ? min(v, red[w][tid]) : max(v, red[w][tid]);
        cbox[6 * static_cast<int64_t>(blockIdx.x) + tid] = v;
    }
}

// A level's node boxes as floats (lo[3], hi[3]) from its chunks' boxes: one
// wave per node over its chunks [blk0, blk0 + nblk).
__global__ __launch_bounds__(64) void k_kdg_box_node(const KdgNode* __restrict__ nodes,
                                                     const unsigned int* __restrict__ cbox,
                                                     float* __restrict__ fbox) {
    const KdgNode N = nodes[blockIdx.x];
    const int lane = threadIdx.x;
    unsigned int lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
    for (uint32_t c = N.blk0 + lane; c < N.blk0 + N.nblk; c += 64) {
        const unsigned int* b = cbox + 6 * static_cast<int64_t>(c);
        for (int d = 0; d < 3; ++d) {
            lo[d] = min(lo[d], b[d]);
            hi[d] = max(hi[d], b[3 + d]);
        }
    }
    for (int d = 0; d < 3; ++d) {
        lo[d] = wave_umin(lo[d]);
        hi[d] = wave_umax(hi[d]);
    }
    const unsigned int v = lane == 0 ? lo[0]
                         : (lane == 1 ? lo[1] : (lane == 2 ? lo[2] : (lane == 3 ? hi[0] : (lane == 4 ? hi[1] : hi[2]))));
    if (lane < 6) fbox[6 * static_cast<int64_t>(blockIdx.x) + lane] = ord2f(v);
}

// Global level: each position's bin (KDG_BINS over its node's longest side)
// and the chunk's cumulative histogram of them (entry b: positions in bins
// 0..b).
__global__ __launch_bounds__(KDG_THREADS) void k_kdg_bin(
    const KdgSet in, const KdgChunk* __restrict__ chunks, const KdgNode* __restrict__ nodes,
    const float* __restrict__ fbox, uint8_t* __restrict__ q8, uint32_t* __restrict__ bh) {
    __shared__ unsigned int h[KDG_BINS];
    const int tid = threadIdx.x;
    const KdgChunk C = chunks[blockIdx.x];
    const uint32_t cnt = nodes[C.node].count;
    h[tid] = 0u;
    __syncthreads();
    if (cnt > TILE) {
        float lo[3], ext[3];
        const int ax = kd_axis(fbox + 6 * static_cast<int64_t>(C.node), lo, ext);
        const float sc = ext[ax] > 0.f ? static_cast<float>(KDG_BINS) / ext[ax] : 0.f;
        const float* __restrict__ c = ax == 0 ? in.x : (ax == 1 ? in.y : in.z);
#pragma unroll
        for (int r = 0; r < KDG_PT; ++r) {
            const uint32_t i = C.begin + tid + r * KDG_THREADS;
            if (i < C.end) {
                const float t = (c[i] - lo[ax]) * sc;
                const int q = min(KDG_BINS - 1, max(0, static_cast<int>(t)));
                q8[i] = static_cast<uint8_t>(q);
                atomicAdd(&h[q], 1u);
            }
        }
    }
    __syncthreads();
    // inclusive scan over the bins (one bin per thread)
    __shared__ uint32_t wsum[KDG_THREADS / 64];
    const int lane = tid & 63, wave = tid >> 6;
    uint32_t inc = h[tid];
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(inc), o, 64));
        if (lane >= o) inc += y;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    for (int w = 0; w < wave; ++w) inc += wsum[w];
    bh[static_cast<int64_t>(blockIdx.x) * KDG_BINS + tid] = inc;
}

// Global level: the nodes' cumulative histograms, summed over their chunks
// in slices (slice s of a node takes its chunks s, s + S, ...).
__global__ __launch_bounds__(KDG_THREADS) void k_kdg_total(
    const KdgNode* __restrict__ nodes, const uint32_t* __restrict__ bh, int S,
    uint32_t* __restrict__ part) {
    const int s = blockIdx.x, tid = threadIdx.x;
    const int64_t h = blockIdx.y;
    const KdgNode N = nodes[h];
    uint32_t tot = 0;
    if (N.count > static_cast<uint32_t>(TILE))
        for (uint32_t b = s; b < N.nblk; b += S) tot += bh[static_cast<int64_t>(N.blk0 + b) * KDG_BINS + tid];
    part[(h * S + s) * KDG_BINS + tid] = tot;
}

// Global level, one workgroup per node: the bin holding the cut (from the
// S slice sums of the cumulative histograms), then each of the node's chunks'
// counts below the cut bin and in it (two entries of its cumulative
// histogram) and their exclusive prefix over the node's chunks.  With
// fbox_next the children's boxes follow from the parent's: cut at the cut
// bin's edges along the split side (a node that does not split hands its box
// to both).
__global__ __launch_bounds__(KDG_THREADS) void k_kdg_select(
    const KdgNode* __restrict__ nodes, const uint32_t* __restrict__ part, int S,
    const uint32_t* __restrict__ bh, const float* __restrict__ fbox, float* __restrict__ fbox_next,
    int2* __restrict__ sel, uint2* __restrict__ boff) {
    __shared__ uint32_t cum[KDG_BINS];
    __shared__ unsigned long long wsum[KDG_THREADS / 64];
    __shared__ int s_bin, s_tie;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t h = blockIdx.x;
    const KdgNode N = nodes[h];
    float b6[6];
    for (int d = 0; d < 6; ++d) b6[d] = fbox[6 * h + d];
    if (N.count <= static_cast<uint32_t>(TILE)) {
        if (fbox_next && tid < 12) fbox_next[12 * h + tid] = b6[tid % 6];
        return;
    }
    uint32_t tot = 0;
    for (int s = 0; s < S; ++s) tot += part[(h * S + s) * KDG_BINS + tid];
    cum[tid] = tot;
    __syncthreads();
    const uint32_t exc = tid > 0 ? cum[tid - 1] : 0u;
    if (exc < N.need && tot >= N.need) {             // exactly one bin
        s_bin = tid;
        s_tie = static_cast<int>(N.need - exc);
    }
    __syncthreads();
    const int bs = s_bin;
    if (tid == 0) sel[h] = make_int2(bs, s_tie);
    if (fbox_next && tid < 2) {
        float lo[3], ext[3];
        const int ax = kd_axis(b6, lo, ext);
        const float w = ext[ax] / static_cast<float>(KDG_BINS);
        float c6[6];
        for (int d = 0; d < 6; ++d) c6[d] = b6[d];
        if (ext[ax] > 0.f) {
            if (tid == 0) c6[3 + ax] = fminf(b6[3 + ax], lo[ax] + static_cast<float>(bs + 1) * w);
            else c6[ax] = fmaxf(b6[ax], lo[ax] + static_cast<float>(bs) * w);
        }
        for (int d = 0; d < 6; ++d) fbox_next[12 * h + 6 * tid + d] = c6[d];
    }
    // the chunks' (below, in) counts, packed high / low, scanned over the node
    unsigned long long carry = 0ull;
    for (uint32_t base = 0; base < N.nblk; base += KDG_THREADS) {
        const uint32_t b = base + tid;
        unsigned long long v = 0ull;
        if (b < N.nblk) {
            const uint32_t* c = bh + static_cast<int64_t>(N.blk0 + b) * KDG_BINS;
            const uint32_t below = bs > 0 ? c[bs - 1] : 0u;
            v = (static_cast<unsigned long long>(below) << 32) | (c[bs] - below);
        }
        unsigned long long inc = v;
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned long long y = __shfl_up(inc, o, 64);
            if (lane >= o) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        unsigned long long before = carry;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        const unsigned long long ex = before + inc - v;
        if (b < N.nblk)
            boff[N.blk0 + b] = make_uint2(static_cast<uint32_t>(ex >> 32), static_cast<uint32_t>(ex));
        for (int w = 0; w < KDG_THREADS / 64; ++w) carry += wsum[w];
        __syncthreads();                              // wsum is rewritten next round
    }
}

// Global level: the stable partition of each chunk into its node's halves
// (a node that does not split is copied).  Wave w takes the chunk's
// positions 512 w + 64 r + lane (r = 0..7), so neighbouring lanes hold
// neighbouring positions and mostly write neighbouring slots; the counts
// before a position come from ballots (within a step), running sums over the
// steps and the waves before it (through LDS).
__global__ __launch_bounds__(KDG_THREADS) void k_kdg_scatter(
    const KdgChunk* __restrict__ chunks, const KdgNode* __restrict__ nodes,
    const uint8_t* __restrict__ q8, const KdgSet in, const int2* __restrict__ sel,
    const uint2* __restrict__ boff, const KdgSet out) {
    static_assert(KDG_CHUNK == 4 * 64 * KDG_PT && KDG_THREADS == 256, "four waves of 8 steps");
    __shared__ uint32_t wb[4], wt[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const KdgChunk C = chunks[blockIdx.x];
    const KdgNode N = nodes[C.node];
    const uint32_t e0 = C.begin + 512 * wave + lane;
    auto move = [&](uint32_t from, int64_t to) {
        out.x[to] = in.x[from];
        out.y[to] = in.y[from];
        out.z[to] = in.z[from];
        out.idx[to] = in.idx[from];
    };
    if (N.count <= static_cast<uint32_t>(TILE)) {
#pragma unroll
        for (int r = 0; r < KDG_PT; ++r)
            if (e0 + 64 * r < C.end) move(e0 + 64 * r, e0 + 64 * r);
        return;
    }
    const int2 sl = sel[C.node];
    const uint2 bo = boff[blockIdx.x];
    auto below_mask = [](uint64_t m) {             // set lanes below this one
        return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                         __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
    };
    bool lb[KDG_PT], tb[KDG_PT];
    uint32_t nb = 0, nt = 0;
#pragma unroll
    for (int r = 0; r < KDG_PT; ++r) {
        const uint32_t e = e0 + 64 * r;
        const int q = e < C.end ? static_cast<int>(q8[e]) : 256;
        lb[r] = q < sl.x;
        tb[r] = q == sl.x;
        nb += __popcll(__ballot(lb[r]));
        nt += __popcll(__ballot(tb[r]));
    }
    if (lane == 0) {
        wb[wave] = nb;
        wt[wave] = nt;
    }
    __syncthreads();
    uint32_t runb = bo.x, runt = bo.y;
    for (int w = 0; w < wave; ++w) {
        runb += wb[w];
        runt += wt[w];
    }
#pragma unroll
    for (int r = 0; r < KDG_PT; ++r) {
        const uint32_t e = e0 + 64 * r;
        const uint64_t mb = __ballot(lb[r]), mt = __ballot(tb[r]);
        const uint32_t before_b = runb + below_mask(mb), before_t = runt + below_mask(mt);
        if (e < C.end) {
            const uint32_t before = before_b + min(before_t, static_cast<uint32_t>(sl.y));
            const bool left = lb[r] || (tb[r] && before_t < static_cast<uint32_t>(sl.y));
            const int64_t dst = left ? static_cast<int64_t>(N.begin) + before
                                     : static_cast<int64_t>(N.begin) + N.need +
                                           (static_cast<int64_t>(e) - N.begin - before);
            // inside the node by construction; the guard keeps a logic error
            // from writing elsewhere
            if (dst >= N.begin && dst < static_cast<int64_t>(N.begin) + N.count) move(e, dst);
        }
        runb += __popcll(mb);
        runt += __popcll(mt);
    }
}

struct SpeciesSplit {
    int n_species;
    int64_t atom_start[MAX_SPECIES + 1];
    int tile_start[MAX_SPECIES + 1];
};

// Grouping by species (stable) as a counting sort straight into the global
// levels' first set: per block of KDG_CHUNK atoms its species counts
// (k_species_count), their exclusive prefix per species over the blocks from
// the species' first position (k_species_prefix), and the scatter of each
// atom's coordinates and index (k_species_scatter; wave w takes the block's
// atoms 512 w + 64 r + lane, ranks within a step from ballots).
__global__ __launch_bounds__(KDG_THREADS) void k_species_count(const uint8_t* __restrict__ spc,
                                                               int64_t n, uint32_t* __restrict__ cnt) {
    __shared__ uint32_t h[MAX_SPECIES];
    const int tid = threadIdx.x;
    if (tid < MAX_SPECIES) h[tid] = 0u;
    __syncthreads();
    const int64_t b0 = static_cast<int64_t>(blockIdx.x) * KDG_CHUNK;
#pragma unroll
    for (int r = 0; r < KDG_PT; ++r) {
        const int64_t i = b0 + tid + r * KDG_THREADS;
        if (i < n) atomicAdd(&h[spc[i] & 0xF], 1u);
    }
    __syncthreads();
    if (tid < MAX_SPECIES) cnt[static_cast<int64_t>(blockIdx.x) * MAX_SPECIES + tid] = h[tid];
}

// one workgroup per species
__global__ __launch_bounds__(KDG_THREADS) void k_species_prefix(const uint32_t* __restrict__ cnt,
                                                                int nblocks, SpeciesSplit split,
                                                                uint32_t* __restrict__ off) {
    __shared__ uint32_t wsum[KDG_THREADS / 64];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int sp = blockIdx.x;
    uint32_t carry = static_cast<uint32_t>(split.atom_start[sp]);
    for (int base = 0; base < nblocks; base += KDG_THREADS) {
        const int b = base + tid;
        const uint32_t v = b < nblocks ? cnt[static_cast<int64_t>(b) * MAX_SPECIES + sp] : 0u;
        uint32_t inc = v;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(inc), o, 64));
            if (lane >= o) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        uint32_t before = carry;
        for (int w = 0; w < wave; ++w) before += wsum[w];
        if (b < nblocks) off[static_cast<int64_t>(b) * MAX_SPECIES + sp] = before + inc - v;
        for (int w = 0; w < KDG_THREADS / 64; ++w) carry += wsum[w];
        __syncthreads();                              // wsum is rewritten next round
    }
}

__global__ __launch_bounds__(KDG_THREADS) void k_species_scatter(
    const float* __restrict__ pos, const uint8_t* __restrict__ spc, int64_t n, int n_species,
    const uint32_t* __restrict__ off, KdgSet out) {
    static_assert(KDG_CHUNK == 4 * 64 * KDG_PT && KDG_THREADS == 256, "four waves of 8 steps");
    __shared__ uint32_t wcnt[4][MAX_SPECIES];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t e0 = static_cast<int64_t>(blockIdx.x) * KDG_CHUNK + 512 * wave + lane;
    auto below_mask = [](uint64_t m) {             // set lanes below this one
        return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                         __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
    };
    int sp[KDG_PT];
#pragma unroll
    for (int r = 0; r < KDG_PT; ++r) {
        const int64_t e = e0 + 64 * r;
        sp[r] = e < n ? static_cast<int>(spc[e] & 0xF) : MAX_SPECIES;
    }
    for (int s = 0; s < n_species; ++s) {
        uint32_t c = 0;
#pragma unroll
        for (int r = 0; r < KDG_PT; ++r) c += __popcll(__ballot(sp[r] == s));
        if (lane == 0) wcnt[wave][s] = c;
    }
    __syncthreads();
    // this wave's first slot per species: the block's offset and the waves before it
    uint32_t base[MAX_SPECIES];
#pragma unroll
    for (int s = 0; s < MAX_SPECIES; ++s) {
        base[s] = 0u;
        if (s < n_species) {
            base[s] = off[static_cast<int64_t>(blockIdx.x) * MAX_SPECIES + s];
            for (int w = 0; w < wave; ++w) base[s] += wcnt[w][s];
        }
    }
#pragma unroll
    for (int r = 0; r < KDG_PT; ++r) {
        const int64_t e = e0 + 64 * r;
        uint32_t dst = 0u;
#pragma unroll
        for (int s = 0; s < MAX_SPECIES; ++s) {
            if (s >= n_species) break;
            const uint64_t m = __ballot(sp[r] == s);
            if (sp[r] == s) dst = base[s] + below_mask(m);
            base[s] += __popcll(m);
        }
        // (codes that disagree with the host's species counts must not
        // write outside the arrays)
        if (e < n && dst < static_cast<uint64_t>(n)) {
            out.x[dst] = pos[3 * e];
            out.y[dst] = pos[3 * e + 1];
            out.z[dst] = pos[3 * e + 2];
            out.idx[dst] = static_cast<uint32_t>(e);
        }
    }
}

// Origin of a unit of 2 or 4 consecutive tiles: the mean of their origins,
// taken pairwise (the moment sum and the staging both use this form).
__device__ __forceinline__ double unit_origin(int n, double a, double b, double c, double d) {
    const double ab = 0.5 * (a + b);
    return n == 2 ? ab : 0.5 * (ab + 0.5 * (c + d));
}

// Record of one atom of a tile: FP32 offset from the tile origin o [m] in
// wavelengths, amplitude as magnitude and phase in turns.  r is the position
// in metres (the float position in angstrom over 1e10).
__device__ __forceinline__ AtomRec atom_record(const double (&r)[3], const double (&o)[3],
                                               double K, uint32_t idx,
                                               const float2* __restrict__ amp) {
    AtomRec rec;
    rec.sx = static_cast<float>((r[0] - o[0]) * K);
    rec.sy = static_cast<float>((r[1] - o[1]) * K);
    rec.sz = static_cast<float>((r[2] - o[2]) * K);
    rec.s2 = rec.sx * rec.sx + rec.sy * rec.sy + rec.sz * rec.sz;
    rec.pad0 = __uint_as_float(idx);     // original atom index (adjoint scatter)
    if (amp) {
        const float2 w = amp[idx];
        rec.mag = sqrtf(w.x * w.x + w.y * w.y);
        rec.ph = atan2f(w.y, w.x) * 0.15915494309189535f;
    } else {
        rec.mag = 1.f;
        rec.ph = 0.f;
    }
    rec.rmag = sqrtf(rec.mag);
    return rec;
}

// One workgroup per tile: FP64 origin at the tile's bounding-box centre, FP32
// offsets in wavelengths, amplitudes; padding lanes get zero weight.
__global__ __launch_bounds__(TILE) void k_tiles(
    const float* __restrict__ pos, const float2* __restrict__ amp,
    const uint32_t* __restrict__ order, SpeciesSplit split, double K,
    TileRec* __restrict__ tiles, AtomRec* __restrict__ atoms,
    unsigned int* __restrict__ radius_bits) {
    __shared__ double lds[TILE / 64];
    const int t = blockIdx.x;
    int s = 0;
    while (s + 1 < split.n_species && t >= split.tile_start[s + 1]) ++s;
    const int64_t begin = split.atom_start[s] + static_cast<int64_t>(t - split.tile_start[s]) * TILE;
    const int64_t end = split.atom_start[s + 1];
    const int count = static_cast<int>(min<int64_t>(TILE, end - begin));
    const int i = threadIdx.x;
    const bool live = i < count;
    uint32_t idx = live ? order[begin + i] : 0u;
    double r[3] = {0.0, 0.0, 0.0};
    if (live)
        for (int d = 0; d < 3; ++d) r[d] = static_cast<double>(pos[3 * (int64_t)idx + d]) / 1e10;
    double o[3];
    for (int d = 0; d < 3; ++d) {
        const double a = block_min(live ? r[d] : 1e300, lds);
        const double b = block_max(live ? r[d] : -1e300, lds);
        o[d] = 0.5 * (a + b);
    }
    AtomRec rec = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float rad = 0.f;
    if (live) {
        rec = atom_record(r, o, K, idx, amp);
        rad = sqrtf(rec.s2);
    }
    atoms[static_cast<int64_t>(t) * TILE + i] = rec;
    const double rmax = block_max(static_cast<double>(rad), lds);
    if (i == 0) {
        TileRec tr;
        tr.ox = o[0];
        tr.oy = o[1];
        tr.oz = o[2];
        tr.species = s;
        tr.count = count;
        tiles[t] = tr;
        atomicMax(radius_bits, __float_as_uint(static_cast<float>(rmax)));
        const float on = static_cast<float>(sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]));
        atomicMax(radius_bits + 1, __float_as_uint(on * 1.0000002f));
    }
}

// LDS index of the slot of position p in k_kd_local: one pad word per 8, so
// that a wave touching the 8 consecutive positions of each lane (lane stride
// 9 words) meets 64 different banks.
__device__ __forceinline__ int kidx(int p) { return p + (p >> 3); }

constexpr int KDL_BINS = 256;      // coordinate bins of an in-LDS split
constexpr int KDL_SUBMAX = 1 << (KDL_LEVELS - 1);   // sub-nodes of the deepest split level

// After a restage (new positions, same k-d order): the radii of the pairs and
// quads of tiles about their unit origins, from the new records.  k_unit_init
// clears the heads (0) and marks the rest (+inf); k_unit_radii, one
// workgroup per tile, takes the largest |offset + tile shift| of its atoms
// into its pair and quad heads (atomic max on the float bits, non-negative).
__global__ void k_unit_init(const uint32_t* __restrict__ tile_pos, int n_tiles,
                            float* __restrict__ grp_r) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tiles) return;
    const int j = tile_pos[t] & 0xFF, ntl = tile_pos[t] >> 8;
    grp_r[2 * static_cast<int64_t>(t)] = ((j & 1) == 0 && j + 1 < ntl) ? 0.f : INFINITY;
    grp_r[2 * static_cast<int64_t>(t) + 1] = ((j & 3) == 0 && j + 3 < ntl) ? 0.f : INFINITY;
}

__global__ __launch_bounds__(TILE) void k_unit_radii(const TileRec* __restrict__ tiles,
                                                     const AtomRec* __restrict__ atoms,
                                                     const uint32_t* __restrict__ tile_pos,
                                                     double K, float* __restrict__ grp_r) {
    __shared__ float red[TILE / 64][2];
    const int t = blockIdx.x, i = threadIdx.x, lane = i & 63, wave = i >> 6;
    const int j = tile_pos[t] & 0xFF, ntl = tile_pos[t] >> 8;
    const int hp = t - (j & 1), hq = t - (j & 3);
    const bool pair = (j & ~1) + 1 < ntl, quad = (j & ~3) + 3 < ntl;
    if (!pair && !quad) return;                      // the whole workgroup
    const TileRec tr = tiles[t];
    float dp[3] = {0.f, 0.f, 0.f}, dq[3] = {0.f, 0.f, 0.f};
    if (pair) {
        const TileRec a = tiles[hp], b = tiles[hp + 1];
        dp[0] = static_cast<float>((tr.ox - unit_origin(2, a.ox, b.ox, 0.0, 0.0)) * K);
        dp[1] = static_cast<float>((tr.oy - unit_origin(2, a.oy, b.oy, 0.0, 0.0)) * K);
        dp[2] = static_cast<float>((tr.oz - unit_origin(2, a.oz, b.oz, 0.0, 0.0)) * K);
    }
    if (quad) {
        const TileRec a = tiles[hq], b = tiles[hq + 1], c = tiles[hq + 2], d = tiles[hq + 3];
        dq[0] = static_cast<float>((tr.ox - unit_origin(4, a.ox, b.ox, c.ox, d.ox)) * K);
        dq[1] = static_cast<float>((tr.oy - unit_origin(4, a.oy, b.oy, c.oy, d.oy)) * K);
        dq[2] = static_cast<float>((tr.oz - unit_origin(4, a.oz, b.oz, c.oz, d.oz)) * K);
    }
    float rp = 0.f, rq = 0.f;
    if (i < tr.count) {
        const AtomRec r = atoms[static_cast<int64_t>(t) * TILE + i];
        const float px = r.sx + dp[0], py = r.sy + dp[1], pz = r.sz + dp[2];
        const float qx = r.sx + dq[0], qy = r.sy + dq[1], qz = r.sz + dq[2];
        rp = sqrtf(px * px + py * py + pz * pz);
        rq = sqrtf(qx * qx + qy * qy + qz * qz);
    }
    for (int o = 32; o > 0; o >>= 1) {
        rp = fmaxf(rp, __shfl_xor(rp, o, 64));
        rq = fmaxf(rq, __shfl_xor(rq, o, 64));
    }
    if (lane == 0) { red[wave][0] = rp; red[wave][1] = rq; }
    __syncthreads();
    if (i == 0) {
        for (int w = 1; w < TILE / 64; ++w) {
            rp = fmaxf(rp, red[w][0]);
            rq = fmaxf(rq, red[w][1]);
        }
        if (pair) atomicMax(reinterpret_cast<unsigned int*>(grp_r) + 2 * static_cast<int64_t>(hp),
                            __float_as_uint(rp));
        if (quad) atomicMax(reinterpret_cast<unsigned int*>(grp_r) + 2 * static_cast<int64_t>(hq) + 1,
                            __float_as_uint(rq));
    }
}

// The last k-d levels and the tile records of one node, in one workgroup of
// 32 << LV threads (8 positions each).  The node's positions
// [begin, begin + count) of the order (at most TILE << LV atoms, a whole
// number of tiles except at a species end) sit in LDS as coordinates by load
// slot and a slot per position.  Each of LV levels splits every sub-node of
// more than one tile by the count rule of the global levels (the left child
// keeps ceil(n / 2T) T positions): a bounding box, the longest side, a
// histogram of the coordinate over that side in KDL_BINS bins, the bin
// holding the cut, and a stable partition (a packed block-wide scan of "below
// the cut bin" and "in the cut bin" flags) in which the first positions of
// the cut bin go left until the count is met.  The halves overlap by at most
// one bin.  Then the node's order is written back and its tiles are emitted
// as k_tiles emits them (the coordinates and atom indices read from the
// global levels' arrays; the order written to order, which may be their
// index array).  Thread tid holds positions 8 tid .. 8 tid + 7, so
// a half-wave covers one tile, which never straddles two sub-nodes.
template <int LV>
__global__ __launch_bounds__(32 << LV) void k_kd_local(
    const KdgSet in, const float2* __restrict__ amp, uint32_t* order,
    const KdNode* __restrict__ nodes, double K, TileRec* __restrict__ tiles,
    AtomRec* __restrict__ atoms, float* __restrict__ grp_r, uint32_t* __restrict__ tile_pos,
    unsigned int* __restrict__ radius_bits) {
    constexpr int THREADS = 32 << LV, MAXN = TILE << LV, PT = MAXN / THREADS;
    constexpr int SUB = (1 << LV) - 1, SUBMAX = 1 << (LV - 1), WAVES = THREADS / 64;
    constexpr int NTILE = 1 << LV;
    static_assert(PT == 8 && WAVES >= SUBMAX, "one wave per deepest sub-node, 8 positions a lane");
    __shared__ float xyz[3][MAXN];                    // coordinates by load slot
    __shared__ uint32_t slot[MAXN + MAXN / 8];        // load slot at kidx(position)
    __shared__ unsigned int hist[SUBMAX][KDL_BINS];
    __shared__ int sub_b[SUB], sub_n[SUB];            // sub-node ranges, level-major heap
    __shared__ int sub_any[LV];                       // some sub-node of the level splits
    __shared__ unsigned int box[SUBMAX][6];
    __shared__ int sel_bin[SUBMAX], sel_tie[SUBMAX];
    __shared__ uint32_t wsum[WAVES];
    __shared__ uint32_t spref[SUBMAX];
    __shared__ double tile_o[NTILE][3];
    __shared__ float tile_r[NTILE], tile_on[NTILE];
    __shared__ float quad_r[WAVES];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const KdNode nd = nodes[blockIdx.x];
    const int n = static_cast<int>(nd.count);
    const int64_t begin = nd.begin;
    const int p0 = PT * tid;                          // this thread's first position

#pragma unroll
    for (int i = 0; i < PT; ++i) {
        const int p = tid + i * THREADS;
        if (p < n) {
            xyz[0][p] = in.x[begin + p];
            xyz[1][p] = in.y[begin + p];
            xyz[2][p] = in.z[begin + p];
        }
        slot[kidx(p)] = static_cast<uint32_t>(p);
    }
    // sub-node (f, j) is heap entry 2^f - 1 + j; its range follows the cuts
    // along the path from the root
    if (tid < SUB) {
        int f = 0;
        while ((2 << f) <= tid + 1) ++f;
        const int j = tid + 1 - (1 << f);
        int b = 0, m = n;
        for (int g = 0; g < f; ++g) {
            const int cut = m > TILE ? (m + 2 * TILE - 1) / (2 * TILE) * TILE : m;
            if ((j >> (f - 1 - g)) & 1) {
                b += cut;
                m -= cut;
            } else {
                m = cut;
            }
        }
        sub_b[tid] = b;
        sub_n[tid] = m;
    }
    if (tid < LV) sub_any[tid] = 0;
    __syncthreads();
    if (tid < SUB && sub_n[tid] > TILE) {
        int f = 0;
        while ((2 << f) <= tid + 1) ++f;
        atomicOr(&sub_any[f], 1);
    }
    __syncthreads();

    for (int f = 0; f < LV; ++f) {
        if (!sub_any[f]) break;                       // no sub-node splits any more
        const int base = (1 << f) - 1, nsub = 1 << f;
        if (tid < 6 * nsub) box[tid / 6][tid % 6] = (tid % 6 < 3) ? 0xFFFFFFFFu : 0u;
        for (int i = tid; i < nsub * KDL_BINS; i += THREADS) hist[i / KDL_BINS][i % KDL_BINS] = 0u;
        int j = 0;
        while (j + 1 < nsub && p0 >= sub_b[base + j + 1]) ++j;
        const int bj = sub_b[base + j], nj = sub_n[base + j];
        const bool splits = p0 < n && nj > TILE;
        uint32_t s[PT];
        int dg[PT];
        unsigned int lo3[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi3[3] = {0u, 0u, 0u};
#pragma unroll
        for (int r = 0; r < PT; ++r) {
            s[r] = slot[kidx(p0 + r)];
            dg[r] = 0;
            if (splits && p0 + r < n)
                for (int d = 0; d < 3; ++d) {
                    const unsigned int o = f2ord(xyz[d][s[r]]);
                    lo3[d] = min(lo3[d], o);
                    hi3[d] = max(hi3[d], o);
                }
        }
        // bounding boxes over the half-wave (one tile of one sub-node)
        for (int d = 0; d < 3; ++d)
            for (int o = 16; o > 0; o >>= 1) {
                lo3[d] = min(lo3[d], static_cast<unsigned int>(__shfl_xor(static_cast<int>(lo3[d]), o, 64)));
                hi3[d] = max(hi3[d], static_cast<unsigned int>(__shfl_xor(static_cast<int>(hi3[d]), o, 64)));
            }
        __syncthreads();                              // box and hist cleared
        if ((lane & 31) == 0 && splits)
            for (int d = 0; d < 3; ++d) {
                atomicMin(&box[j][d], lo3[d]);
                atomicMax(&box[j][3 + d], hi3[d]);
            }
        __syncthreads();
        if (splits) {
            float l3[3], e3[3];
            for (int d = 0; d < 3; ++d) {
                l3[d] = ord2f(box[j][d]);
                e3[d] = ord2f(box[j][3 + d]) - l3[d];
            }
            int ax = 0;
            if (e3[1] > e3[ax]) ax = 1;
            if (e3[2] > e3[ax]) ax = 2;
            const float sc = e3[ax] > 0.f ? static_cast<float>(KDL_BINS) / e3[ax] : 0.f;
#pragma unroll
            for (int r = 0; r < PT; ++r)
                if (p0 + r < n) {
                    const float t = (xyz[ax][s[r]] - l3[ax]) * sc;
                    dg[r] = min(KDL_BINS - 1, max(0, static_cast<int>(t)));
                    atomicAdd(&hist[j][dg[r]], 1u);
                }
        }
        __syncthreads();
        // the bin holding the cut, one wave per sub-node
        if (wave < nsub && sub_n[base + wave] > TILE) {
            const int mm = sub_n[base + wave];
            const uint32_t need = static_cast<uint32_t>((mm + 2 * TILE - 1) / (2 * TILE) * TILE);
            uint32_t c4[4], own = 0;
            for (int q = 0; q < 4; ++q) {
                c4[q] = hist[wave][4 * lane + q];
                own += c4[q];
            }
            uint32_t inc = own;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(inc), o, 64));
                if (lane >= o) inc += y;
            }
            const unsigned long long ball = __ballot(inc >= need);
            if (lane == __ffsll(ball) - 1) {
                uint32_t below = inc - own;
                int q = 0;
                while (q < 3 && below + c4[q] < need) below += c4[q++];
                sel_bin[wave] = 4 * lane + q;
                sel_tie[wave] = static_cast<int>(need - below);
            }
        }
        __syncthreads();
        // packed counts (below the cut bin) << 16 | (in the cut bin), scanned
        // over the positions
        const int cb = splits ? sel_bin[j] : -1;
        uint32_t fl[PT], own = 0;
#pragma unroll
        for (int r = 0; r < PT; ++r) {
            fl[r] = 0u;
            if (splits && p0 + r < n) fl[r] = dg[r] < cb ? 0x10000u : (dg[r] == cb ? 1u : 0u);
            own += fl[r];
        }
        uint32_t inc = own;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(inc), o, 64));
            if (lane >= o) inc += y;
        }
        if (lane == 63) wsum[wave] = inc;
        __syncthreads();
        uint32_t excl = inc - own;
        for (int w = 0; w < wave; ++w) excl += wsum[w];
        if (splits && p0 == bj) spref[j] = excl;
        __syncthreads();
        // stable partition; every slot of the level was read before the
        // barriers above
        if (splits) {
            const int tie = sel_tie[j];
            const int need = (nj + 2 * TILE - 1) / (2 * TILE) * TILE;
            uint32_t run = excl - spref[j];
#pragma unroll
            for (int r = 0; r < PT; ++r) {
                const int p = p0 + r;
                if (p < n) {
                    const int lb = static_cast<int>(run >> 16), tb = static_cast<int>(run & 0xFFFFu);
                    const int before = lb + min(tb, tie);
                    const bool left = fl[r] == 0x10000u || (fl[r] == 1u && tb < tie);
                    const int dst = left ? bj + before : bj + need + (p - bj - before);
                    slot[kidx(dst)] = s[r];
                }
                run += fl[r];
            }
        }
        __syncthreads();
    }

    // tile origins (bounding-box centres) and radii, one half-wave per tile;
    // the minimum of float / 1e10 in FP64 is the minimum float / 1e10
    const int ntl = (n + TILE - 1) / TILE;
    {
        uint32_t s[PT];
        float lo3[3] = {INFINITY, INFINITY, INFINITY}, hi3[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
        for (int r = 0; r < PT; ++r) {
            s[r] = slot[kidx(p0 + r)];
            if (p0 + r < n)
                for (int d = 0; d < 3; ++d) {
                    lo3[d] = fminf(lo3[d], xyz[d][s[r]]);
                    hi3[d] = fmaxf(hi3[d], xyz[d][s[r]]);
                }
        }
        for (int d = 0; d < 3; ++d)
            for (int o = 16; o > 0; o >>= 1) {
                lo3[d] = fminf(lo3[d], __shfl_xor(lo3[d], o, 64));
                hi3[d] = fmaxf(hi3[d], __shfl_xor(hi3[d], o, 64));
            }
        double o[3];
        for (int d = 0; d < 3; ++d)
            o[d] = 0.5 * (static_cast<double>(lo3[d]) / 1e10 + static_cast<double>(hi3[d]) / 1e10);
        float rad = 0.f;
#pragma unroll
        for (int r = 0; r < PT; ++r)
            if (p0 + r < n) {
                double rr[3];
                for (int d = 0; d < 3; ++d) rr[d] = static_cast<double>(xyz[d][s[r]]) / 1e10;
                rad = fmaxf(rad, sqrtf(atom_record(rr, o, K, 0u, nullptr).s2));
            }
        for (int q = 16; q > 0; q >>= 1) rad = fmaxf(rad, __shfl_xor(rad, q, 64));
        const int t = p0 / TILE;
        if ((lane & 31) == 0 && p0 < n) {
            for (int d = 0; d < 3; ++d) tile_o[t][d] = o[d];
            tile_r[t] = rad;
            tile_on[t] = static_cast<float>(sqrt(o[0] * o[0] + o[1] * o[1] + o[2] * o[2])) * 1.0000002f;
            TileRec tr;
            tr.ox = o[0];
            tr.oy = o[1];
            tr.oz = o[2];
            tr.species = static_cast<int>(nd.species);
            tr.count = min(TILE, n - t * TILE);
            tiles[nd.tile0 + t] = tr;
            tile_pos[nd.tile0 + t] = static_cast<uint32_t>(t) | (static_cast<uint32_t>(ntl) << 8);
        }
    }
    // radii of the pairs (one wave) and quads (two waves) of tiles about their
    // unit origins; units stay inside the node
    __syncthreads();                                  // tile_o complete
    {
        const int t = p0 / TILE;
        float rp = 0.f, rq = 0.f;
        const int hp = t & ~1, hq = t & ~3;
        const bool pair = hp + 1 < ntl, quad = hq + 3 < ntl;
        double op[3], oq[3];
        for (int d = 0; d < 3; ++d) {
            op[d] = unit_origin(2, tile_o[hp][d], tile_o[min(hp + 1, NTILE - 1)][d], 0.0, 0.0);
            oq[d] = unit_origin(4, tile_o[hq][d], tile_o[min(hq + 1, NTILE - 1)][d],
                                tile_o[min(hq + 2, NTILE - 1)][d], tile_o[min(hq + 3, NTILE - 1)][d]);
        }
        const uint32_t* sp = slot;
#pragma unroll
        for (int r = 0; r < PT; ++r)
            if (p0 + r < n) {
                const uint32_t sl = sp[kidx(p0 + r)];
                double rr[3];
                for (int d = 0; d < 3; ++d) rr[d] = static_cast<double>(xyz[d][sl]) / 1e10;
                if (pair) rp = fmaxf(rp, sqrtf(atom_record(rr, op, K, 0u, nullptr).s2));
                if (quad) rq = fmaxf(rq, sqrtf(atom_record(rr, oq, K, 0u, nullptr).s2));
            }
        for (int q = 32; q > 0; q >>= 1) {
            rp = fmaxf(rp, __shfl_xor(rp, q, 64));
            rq = fmaxf(rq, __shfl_xor(rq, q, 64));
        }
        if (lane == 0) quad_r[wave] = rq;
        __syncthreads();
        if ((lane & 31) == 0 && p0 < n) {
            const bool ph = pair && t == hp, qh = quad && t == hq;
            grp_r[2 * (static_cast<int64_t>(nd.tile0) + t)] = ph ? rp : INFINITY;
            grp_r[2 * (static_cast<int64_t>(nd.tile0) + t) + 1] =
                qh ? fmaxf(quad_r[wave], quad_r[min(wave + 1, WAVES - 1)]) : INFINITY;
        }
    }

    // the node's order; every read of the old order has returned (its value
    // is stored to LDS) before the barrier that precedes the first write
    uint32_t sl[PT], at[PT];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PT; ++i) {
        const int p = tid + i * THREADS;
        sl[i] = 0u;
        at[i] = 0u;
        if (p < n) {
            sl[i] = slot[kidx(p)];
            at[i] = in.idx[begin + sl[i]];
        }
    }
    __syncthreads();                                  // slot reads above are done
#pragma unroll
    for (int i = 0; i < PT; ++i) {
        const int p = tid + i * THREADS;
        if (p < n) slot[kidx(p)] = at[i];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PT; ++i) {
        const int p = tid + i * THREADS;
        if (p < n) order[begin + p] = slot[kidx(p)];
    }
    if (tid == 0) {
        float rm = 0.f, om = 0.f;
        for (int t = 0; t < ntl; ++t) {
            rm = fmaxf(rm, tile_r[t]);
            om = fmaxf(om, tile_on[t]);
        }
        atomicMax(radius_bits, __float_as_uint(rm));
        atomicMax(radius_bits + 1, __float_as_uint(om));
    }
    // atom records, coalesced: position tid + i THREADS
#pragma unroll
    for (int i = 0; i < PT; ++i) {
        const int p = tid + i * THREADS;
        const int t = p / TILE;
        if (t >= ntl) break;
        AtomRec rec = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        if (p < n) {
            double rr[3];
            for (int d = 0; d < 3; ++d) rr[d] = static_cast<double>(xyz[d][sl[i]]) / 1e10;
            const double o[3] = {tile_o[t][0], tile_o[t][1], tile_o[t][2]};
            rec = atom_record(rr, o, K, at[i], amp);
        }
        atoms[(static_cast<int64_t>(nd.tile0) + t) * TILE + (p % TILE)] = rec;
    }
}

// ---------------------------------------------------------------------- sum

struct SumArgs {
    const AtomRec* atoms;
    const TileRec* tiles;
    const FrameRec* frames;
    const double* cx;
    const double* cy;
    const double* cz;
    const double* theta0;
    const double* rho0;
    const double* qcut;
    double2* partial;
    int n_tiles;
    int tiles_per_slice;
    int n_slices;
    int npix;
    double K;         // 1 / lambda
    double k;         // 2 pi / lambda
    double pol_deg;
    double pol_e[3];
    int remove_forward;
    int apply_pol;
    int apply_decay;
    int n_species;
    SpeciesConst spc[MAX_SPECIES];
};

template <int PIX, bool SECOND>
__global__ __launch_bounds__(SUM_THREADS) void k_sum(const SumArgs A) {
    const int frame = blockIdx.z;
    const int slice = blockIdx.y;
    const FrameRec& F = A.frames[frame];
    const int t0 = slice * A.tiles_per_slice;
    const int t1 = min(A.n_tiles, t0 + A.tiles_per_slice);

    int pix[PIX];
    bool valid[PIX];
#pragma unroll
    for (int i = 0; i < PIX; ++i) {
        const int p = blockIdx.x * (SUM_THREADS * PIX) + i * SUM_THREADS + threadIdx.x;
        valid[i] = p < A.npix;
        pix[i] = valid[i] ? p : A.npix - 1;
    }
    double acc_r[PIX], acc_i[PIX];
#pragma unroll
    for (int i = 0; i < PIX; ++i) { acc_r[i] = 0.0; acc_i[i] = 0.0; }

    const double R00 = F.R[0], R01 = F.R[1], R02 = F.R[2];
    const double R10 = F.R[3], R11 = F.R[4], R12 = F.R[5];
    const double R20 = F.R[6], R21 = F.R[7], R22 = F.R[8];
    const float xs0 = static_cast<float>(F.xs[0]);
    const float xs1 = static_cast<float>(F.xs[1]);
    const float xs2 = static_cast<float>(F.xs[2]);

    for (int t = t0; t < t1; ++t) {
        const TileRec tr = A.tiles[t];
        const SpeciesConst& sc = A.spc[tr.species];
        // tile origin in the lab frame
        const double Ox = fma(R00, tr.ox, fma(R01, tr.oy, fma(R02, tr.oz, F.T[0])));
        const double Oy = fma(R10, tr.ox, fma(R11, tr.oy, fma(R12, tr.oz, F.T[1])));
        const double Oz = fma(R20, tr.ox, fma(R21, tr.oy, fma(R22, tr.oz, F.T[2])));
        const double O2 = Ox * Ox + Oy * Oy + Oz * Oz;

        float vx[PIX], vy[PIX], vz[PIX], gr[PIX], gi[PIX], hh[PIX];
#pragma unroll
        for (int i = 0; i < PIX; ++i) {
            const int p = pix[i];
            const double tx = A.cx[p], ty = A.cy[p], tz = A.cz[p];
            const double Dx = tx - Ox, Dy = ty - Oy, Dz = tz - Oz;
            const double rho = sqrt(Dx * Dx + Dy * Dy + Dz * Dz);
            const double r0 = A.rho0[p];
            // rho - rho0 without cancellation: (|O|^2 - 2 T.O) / (rho + rho0)
            const double dR = (O2 - 2.0 * (tx * Ox + ty * Oy + tz * Oz)) / (rho + r0);
            double th = A.theta0[p] + A.K * (Ox + dR);
            th -= rint(th);                          // [-0.5, 0.5]
            const double inv = 1.0 / rho;
            const double ux = Dx * inv, uy = Dy * inv, uz = Dz * inv;
            const double wx = 1.0 - ux, wy = -uy, wz = -uz;  // x_hat - u_hat (lab)
            vx[i] = static_cast<float>(R00 * wx + R10 * wy + R20 * wz);
            vy[i] = static_cast<float>(R01 * wx + R11 * wy + R21 * wz);
            vz[i] = static_cast<float>(R02 * wx + R12 * wy + R22 * wz);
            hh[i] = static_cast<float>(0.5 / (A.K * rho));
            // 1 - cos(2 theta) without cancellation for forward pixels
            const double omc = (Dx > 0.0) ? (Dy * Dy + Dz * Dz) * inv / (rho + Dx) : 1.0 - ux;
            const double Q = A.k * sqrt(fmax(0.0, 2.0 * omc));
            const float sq = static_cast<float>(Q * (0.25e-10 / PI));
            const float ss = sq * sq;
            float f0 = sc.c;
#pragma unroll
            for (int m = 0; m < 5; ++m) f0 += sc.a[m] * __expf(-sc.b[m] * ss);
            float s_re = f0 + sc.fp, s_im = sc.fpp;
            if (A.remove_forward && Q < A.qcut[p]) {
                s_re -= sc.f0z + sc.fp;
                s_im -= sc.fpp;
            }
            double scale = R_E;
            if (A.apply_decay) scale *= r0 * inv;
            if (A.apply_pol) {
                const double er = A.pol_e[0] * ux + A.pol_e[1] * uy + A.pol_e[2] * uz;
                const double dv = ux;
                double P = 0.5 * (1.0 - A.pol_deg) * (1.0 + dv * dv) + A.pol_deg * (1.0 - er * er);
                P = fmin(1.0, fmax(0.0, P));
                scale *= sqrt(P);
            }
            float sn, cs;
            sincospif(2.0f * static_cast<float>(th), &sn, &cs);
            const float scf = static_cast<float>(scale);
            const float br = s_re * scf, bi = s_im * scf;
            gr[i] = br * cs - bi * sn;
            gi[i] = br * sn + bi * cs;
        }

        float S_r[PIX], S_i[PIX];
#pragma unroll
        for (int i = 0; i < PIX; ++i) { S_r[i] = 0.f; S_i[i] = 0.f; }
        const AtomRec* __restrict__ at = A.atoms + static_cast<int64_t>(t) * TILE;
        const int cnt = tr.count;
#pragma unroll 2
        for (int j = 0; j < cnt; ++j) {
            const AtomRec a = at[j];
            float xsd = 0.f;
            if (SECOND) xsd = fmaf(xs2, a.sz, fmaf(xs1, a.sy, xs0 * a.sx));
#pragma unroll
            for (int i = 0; i < PIX; ++i) {
                float psi = fmaf(vz[i], a.sz, fmaf(vy[i], a.sy, vx[i] * a.sx));
                if (SECOND) {
                    const float u = xsd - psi;          // u_hat . sigma (sample frame)
                    psi = fmaf(hh[i], fmaf(-u, u, a.s2), psi);
                }
                const float r = __builtin_amdgcn_fractf(psi + a.ph);
                const float sn = __builtin_amdgcn_sinf(r);
                const float cs = __builtin_amdgcn_cosf(r);
                S_r[i] = fmaf(a.mag, cs, S_r[i]);
                S_i[i] = fmaf(a.mag, sn, S_i[i]);
            }
        }
#pragma unroll
        for (int i = 0; i < PIX; ++i) {
            const double g_r = gr[i], g_i = gi[i], s_r = S_r[i], s_i = S_i[i];
            acc_r[i] = fma(g_r, s_r, fma(-g_i, s_i, acc_r[i]));
            acc_i[i] = fma(g_r, s_i, fma(g_i, s_r, acc_i[i]));
        }
    }
    double2* out = A.partial + (static_cast<int64_t>(frame) * A.n_slices + slice) * A.npix;
#pragma unroll
    for (int i = 0; i < PIX; ++i)
        if (valid[i]) out[pix[i]] = make_double2(acc_r[i], acc_i[i]);
}

// Exact per-pair form of the sum, with no expansion in tile radius / distance:
// the form every other kernel declines to (detectors within ~1e6 tile radii,
// nanometres to millimetres away) and the one that applies the analyser, a
// per-ray acceptance (Beam.py:2350-2378).  The reference's general kernel
// (interaction_kernal, Beam.py:1912-2412) in its exact mode.
//
// Per (tile, pixel), FP64, sample frame: the vector from the tile origin to
// the pixel, D = rho u, and the origin phase
//     Theta = frac(Theta0 + K (X_t + rho - rho0)),
//     rho - rho0 = (|O|^2 - 2 T.O) / (rho + rho0)          (no cancellation).
// Per pair, FP32 and exact, with sigma the atom's offset in wavelengths,
// h = 1 / (K rho), s = u . sigma, x = R^T x_hat and v = x - u:
//     r / rho = q = sqrt(1 + t),  t = h (|sigma|^2 h - 2 s),
//     psi = x . sigma + K (r - rho) = v . sigma + (s t / (1 + q) + |sigma|^2 h) / (1 + q),
//     1 - cos 2theta = (m0 + h psi) / q,  m0 = 1 - x . u   (m0 from FP64),
//     e x r_hat = (e x u - h e x sigma) / q,  R0 / r = (R0 / rho) / q,
// so nothing of size rho enters the FP32 arithmetic and every difference
// that could cancel is formed from small terms.  The form factor (five
// v_exp_f32), sqrt(P), the decay and the analyser weight are evaluated per
// pair (pair_physics.h); the tile sum multiplies r_e e^{2 pi i Theta} into
// the FP64 accumulators.
struct ExactArgs {
    const AtomRec* atoms;
    const TileRec* tiles;
    const FrameRec* frames;
    const double* cx;
    const double* cy;
    const double* cz;
    const double* theta0;
    const double* rho0;
    const double* qcut;
    double2* partial;
    int n_tiles;
    int tiles_per_slice;
    int n_slices;
    int npix;
    double K;         // 1 / lambda
    double k;         // 2 pi / lambda
    double pol_e[3];  // polarization vector (lab)
    double ana_c[3];  // analyser axis (lab, unit)
    float pol_deg;
    float ana_accept; // top-hat half-angle [rad]
    float ana_inv_hw; // 1 / Darwin half-width [1/rad], 0 for none
    int remove_forward;
    int apply_decay;
    FormRow form[MAX_SPECIES];
};

template <int PIX, bool POL, int ANA>
__global__ __launch_bounds__(SUM_THREADS) void k_sum_exact(const ExactArgs A) {
    const int frame = blockIdx.z;
    const int slice = blockIdx.y;
    const FrameRec& F = A.frames[frame];
    const int t0 = slice * A.tiles_per_slice;
    const int t1 = min(A.n_tiles, t0 + A.tiles_per_slice);

    int pix[PIX];
    bool valid[PIX];
    float ocut[PIX];
#pragma unroll
    for (int i = 0; i < PIX; ++i) {
        const int p = blockIdx.x * (SUM_THREADS * PIX) + i * SUM_THREADS + threadIdx.x;
        valid[i] = p < A.npix;
        pix[i] = valid[i] ? p : A.npix - 1;
        // Q < Q_cut  <=>  1 - cos 2theta < (Q_cut / k)^2 / 2
        const double qc = A.qcut[pix[i]] / A.k;
        ocut[i] = A.remove_forward ? static_cast<float>(0.5 * qc * qc) : -1.0f;
    }
    double acc_r[PIX], acc_i[PIX];
#pragma unroll
    for (int i = 0; i < PIX; ++i) { acc_r[i] = 0.0; acc_i[i] = 0.0; }

    const double R00 = F.R[0], R01 = F.R[1], R02 = F.R[2];
    const double R10 = F.R[3], R11 = F.R[4], R12 = F.R[5];
    const double R20 = F.R[6], R21 = F.R[7], R22 = F.R[8];
    // lab vectors in the sample frame: R^T x_hat, R^T e, R^T c
    const double xs0 = F.xs[0], xs1 = F.xs[1], xs2 = F.xs[2];
    const double es0 = R00 * A.pol_e[0] + R10 * A.pol_e[1] + R20 * A.pol_e[2];
    const double es1 = R01 * A.pol_e[0] + R11 * A.pol_e[1] + R21 * A.pol_e[2];
    const double es2 = R02 * A.pol_e[0] + R12 * A.pol_e[1] + R22 * A.pol_e[2];
    const double cs0 = R00 * A.ana_c[0] + R10 * A.ana_c[1] + R20 * A.ana_c[2];
    const double cs1 = R01 * A.ana_c[0] + R11 * A.ana_c[1] + R21 * A.ana_c[2];
    const double cs2 = R02 * A.ana_c[0] + R12 * A.ana_c[1] + R22 * A.ana_c[2];
    const float fes0 = static_cast<float>(es0), fes1 = static_cast<float>(es1),
                fes2 = static_cast<float>(es2);
    const float fcs0 = static_cast<float>(cs0), fcs1 = static_cast<float>(cs1),
                fcs2 = static_cast<float>(cs2);
    const float pdeg = A.pol_deg;

    for (int t = t0; t < t1; ++t) {
        const TileRec tr = A.tiles[t];
        const FormRow& fr = A.form[tr.species];
        const double Ox = fma(R00, tr.ox, fma(R01, tr.oy, fma(R02, tr.oz, F.T[0])));
        const double Oy = fma(R10, tr.ox, fma(R11, tr.oy, fma(R12, tr.oz, F.T[1])));
        const double Oz = fma(R20, tr.ox, fma(R21, tr.oy, fma(R22, tr.oz, F.T[2])));
        const double O2 = Ox * Ox + Oy * Oy + Oz * Oz;

        float ux[PIX], uy[PIX], uz[PIX], vx[PIX], vy[PIX], vz[PIX], hh[PIX], m0[PIX];
        float dmul[PIX], dadd[PIX], gr[PIX], gi[PIX];
        float eux[PIX], euy[PIX], euz[PIX], cux[PIX], cuy[PIX], cuz[PIX], c0[PIX];
#pragma unroll
        for (int i = 0; i < PIX; ++i) {
            const int p = pix[i];
            const double tx = A.cx[p], ty = A.cy[p], tz = A.cz[p];
            const double Dx = tx - Ox, Dy = ty - Oy, Dz = tz - Oz;
            const double rho = sqrt(Dx * Dx + Dy * Dy + Dz * Dz);
            const double r0 = A.rho0[p];
            const double dR = (O2 - 2.0 * (tx * Ox + ty * Oy + tz * Oz)) / (rho + r0);
            double th = A.theta0[p] + A.K * (Ox + dR);
            th -= rint(th);
            const double inv = 1.0 / rho;
            const double lx = Dx * inv, ly = Dy * inv, lz = Dz * inv;   // lab unit vector
            const double sx = R00 * lx + R10 * ly + R20 * lz;           // sample frame
            const double sy = R01 * lx + R11 * ly + R21 * lz;
            const double sz = R02 * lx + R12 * ly + R22 * lz;
            ux[i] = static_cast<float>(sx);
            uy[i] = static_cast<float>(sy);
            uz[i] = static_cast<float>(sz);
            vx[i] = static_cast<float>(xs0 - sx);
            vy[i] = static_cast<float>(xs1 - sy);
            vz[i] = static_cast<float>(xs2 - sz);
            hh[i] = static_cast<float>(inv / A.K);
            m0[i] = static_cast<float>((Dx > 0.0) ? (Dy * Dy + Dz * Dz) * inv / (rho + Dx) : 1.0 - lx);
            dmul[i] = (A.apply_decay && r0 > 0.0) ? static_cast<float>(r0 * inv) : 0.0f;
            dadd[i] = (A.apply_decay && r0 > 0.0) ? 0.0f : 1.0f;
            if (POL) {
                eux[i] = static_cast<float>(es1 * sz - es2 * sy);
                euy[i] = static_cast<float>(es2 * sx - es0 * sz);
                euz[i] = static_cast<float>(es0 * sy - es1 * sx);
            }
            if (ANA) {
                cux[i] = static_cast<float>(cs1 * sz - cs2 * sy);
                cuy[i] = static_cast<float>(cs2 * sx - cs0 * sz);
                cuz[i] = static_cast<float>(cs0 * sy - cs1 * sx);
                c0[i] = static_cast<float>(cs0 * sx + cs1 * sy + cs2 * sz);
            }
            const float thf = static_cast<float>(th);
            gr[i] = static_cast<float>(R_E) * __builtin_amdgcn_cosf(thf);
            gi[i] = static_cast<float>(R_E) * __builtin_amdgcn_sinf(thf);
        }

        float S_r[PIX], S_i[PIX];
#pragma unroll
        for (int i = 0; i < PIX; ++i) { S_r[i] = 0.f; S_i[i] = 0.f; }
        const AtomRec* __restrict__ at = A.atoms + static_cast<int64_t>(t) * TILE;
        const int cnt = tr.count;
        for (int j = 0; j < cnt; ++j) {
            const AtomRec a = at[j];
            // per-atom cross products with the polarization and analyser vectors
            float ex_ = 0.f, ey_ = 0.f, ez_ = 0.f, cx_ = 0.f, cy_ = 0.f, cz_ = 0.f, cd_ = 0.f;
            if (POL) {
                ex_ = fmaf(fes1, a.sz, -fes2 * a.sy);
                ey_ = fmaf(fes2, a.sx, -fes0 * a.sz);
                ez_ = fmaf(fes0, a.sy, -fes1 * a.sx);
            }
            if (ANA) {
                cx_ = fmaf(fcs1, a.sz, -fcs2 * a.sy);
                cy_ = fmaf(fcs2, a.sx, -fcs0 * a.sz);
                cz_ = fmaf(fcs0, a.sy, -fcs1 * a.sx);
                cd_ = fmaf(fcs2, a.sz, fmaf(fcs1, a.sy, fcs0 * a.sx));
            }
#pragma unroll
            for (int i = 0; i < PIX; ++i) {
                const float h = hh[i];
                const float s = fmaf(uz[i], a.sz, fmaf(uy[i], a.sy, ux[i] * a.sx));
                const float s2h = a.s2 * h;
                const float tt = (s2h - 2.0f * s) * h;
                const float q2 = fmaxf(1.0f + tt, 1e-30f);
                const float invq = __builtin_amdgcn_rsqf(q2);
                const float q = q2 * invq;
                const float ip = __builtin_amdgcn_rcpf(1.0f + q);
                const float corr = fmaf(s * tt, ip, s2h) * ip;
                const float psi = fmaf(vz[i], a.sz, fmaf(vy[i], a.sy, fmaf(vx[i], a.sx, corr)));
                const float omc = fmaf(h, psi, m0[i]) * invq;
                float sr = form_f0(fr.a, fr.nb, fr.c, omc) + fr.fp, si = fr.fpp;
                if (omc < ocut[i]) {
                    sr -= fr.f0z + fr.fp;
                    si -= fr.fpp;
                }
                float wt = a.mag * fmaf(dmul[i], invq, dadd[i]);
                if (POL) {
                    const float wx = fmaf(-h, ex_, eux[i]), wy = fmaf(-h, ey_, euy[i]),
                                wz = fmaf(-h, ez_, euz[i]);
                    const float ec2 = fmaf(wz, wz, fmaf(wy, wy, wx * wx)) * (invq * invq);
                    wt *= pol_factor(1.0f - omc, ec2, pdeg);
                }
                if (ANA) {
                    const float wx = fmaf(-h, cx_, cux[i]), wy = fmaf(-h, cy_, cuy[i]),
                                wz = fmaf(-h, cz_, cuz[i]);
                    const float sd = __builtin_sqrtf(fmaf(wz, wz, fmaf(wy, wy, wx * wx)));
                    wt *= analyser_weight<ANA>(sd, fmaf(-h, cd_, c0[i]), A.ana_accept,
                                               A.ana_inv_hw);
                }
                // skip an atom sitting on the pixel (r == 0, Beam.py:765)
                wt = (1.0f + tt) > 0.0f ? wt : 0.0f;
                const float arg = __builtin_amdgcn_fractf(psi + a.ph);
                const float cr = wt * __builtin_amdgcn_cosf(arg);
                const float ci = wt * __builtin_amdgcn_sinf(arg);
                S_r[i] = fmaf(sr, cr, fmaf(-si, ci, S_r[i]));
                S_i[i] = fmaf(sr, ci, fmaf(si, cr, S_i[i]));
            }
        }
#pragma unroll
        for (int i = 0; i < PIX; ++i) {
            const double g_r = gr[i], g_i = gi[i], s_r = S_r[i], s_i = S_i[i];
            acc_r[i] = fma(g_r, s_r, fma(-g_i, s_i, acc_r[i]));
            acc_i[i] = fma(g_r, s_i, fma(g_i, s_r, acc_i[i]));
        }
    }
    double2* out = A.partial + (static_cast<int64_t>(frame) * A.n_slices + slice) * A.npix;
#pragma unroll
    for (int i = 0; i < PIX; ++i)
        if (valid[i]) out[pix[i]] = make_double2(acc_r[i], acc_i[i]);
}

// Recurrence form of the sum, used when every tile origin is close to the lab
// origin compared with the pixel distance (far field).  One lane owns KP
// consecutive pixels of a detector row.  Across those pixels the direction
// v = R^T (x_hat - u_hat) is a straight line in the pixel index to within the
// row curvature (the host bounds it), so per atom
//     e^{2 pi i v_m . sigma} = e^{2 pi i v_0 . sigma} (e^{2 pi i g . sigma})^m
// and one FP32 sincos plus a short Taylor series for the small step phasor
// replace KP sincos pairs.  Per (tile, pixel) the origin phase is the
// far-field expansion
//     Theta = Theta0 + K (X_t - u0 . O + (|O|^2 - (u0 . O)^2) / (2 rho0)),
// exact to third order in |O| / rho0, and the factor G comes from the
// per-pixel table.  With SECOND the atom's own second-order phase
// (|sigma|^2 - (u . sigma)^2) / (2 K rho) is added, evaluated once per lane at
// the middle of its KP pixels (across them it changes by ~1e-10 rev).
struct RecArgs {
    const AtomRec* atoms;
    const TileRec* tiles;
    const FrameRec* frames;
    const double4* u0h;
    const double* theta0;
    const float2* gtab;
    double2* partial;
    int n_tiles;
    int tiles_per_slice;
    int n_slices;
    int Ny;
    int Nz;
    int npix;
    int nbx;          // lane blocks per detector row
    double K;         // 1 / lambda
};

template <int KP, bool SECOND>
__global__ __launch_bounds__(SUM_THREADS) void k_sum_rec(const RecArgs A) {
    const int frame = blockIdx.z;
    const int slice = blockIdx.y;
    const FrameRec& F = A.frames[frame];
    const int t0 = slice * A.tiles_per_slice;
    const int t1 = min(A.n_tiles, t0 + A.tiles_per_slice);
    const int lanes = A.Nz * A.nbx;
    const int lb = blockIdx.x * SUM_THREADS + threadIdx.x;
    const bool lane_ok = lb < lanes;
    const int lbc = lane_ok ? lb : 0;
    const int iz = lbc / A.nbx;
    const int iy0 = (lbc % A.nbx) * KP;
    const int nv = min(KP, A.Ny - iy0);           // valid pixels of this lane (>= 1)
    const int pbase = iz * A.Ny + iy0;

    // FP64 accumulators live in LDS: the atom loop keeps only the FP32 tile
    // sums in registers, which doubles the waves a SIMD can hold.
    __shared__ double2 acc[KP][SUM_THREADS];
#pragma unroll
    for (int m = 0; m < KP; ++m) acc[m][threadIdx.x] = make_double2(0.0, 0.0);

    const double R00 = F.R[0], R01 = F.R[1], R02 = F.R[2];
    const double R10 = F.R[3], R11 = F.R[4], R12 = F.R[5];
    const double R20 = F.R[6], R21 = F.R[7], R22 = F.R[8];
    const int m_last = nv - 1, m_a = (nv - 1) / 2, m_b = nv / 2;
    const float xs0 = static_cast<float>(F.xs[0]);
    const float xs1 = static_cast<float>(F.xs[1]);
    const float xs2 = static_cast<float>(F.xs[2]);
    const float hmid = 0.5f * static_cast<float>(m_last);

    for (int t = t0; t < t1; ++t) {
        const TileRec tr = A.tiles[t];
        // Re-read the per-pixel tables every tile instead of holding 40 B per
        // pixel in registers for the whole loop: tying the index to the tile
        // record (count <= 256, so the shift is 0) keeps the loads inside the
        // loop, and occupancy is worth more than the reloads.
        const int pb = pbase + (tr.count >> 16);
        const double Ox = fma(R00, tr.ox, fma(R01, tr.oy, fma(R02, tr.oz, F.T[0])));
        const double Oy = fma(R10, tr.ox, fma(R11, tr.oy, fma(R12, tr.oz, F.T[1])));
        const double Oz = fma(R20, tr.ox, fma(R21, tr.oy, fma(R22, tr.oz, F.T[2])));
        const double O2 = Ox * Ox + Oy * Oy + Oz * Oz;

        // v in the sample frame at pixel offset m of this lane
        auto vdir = [&](int m, double& sx, double& sy, double& sz) {
            const double4 u = A.u0h[pb + m];
            const double uo = u.x * Ox + u.y * Oy + u.z * Oz;
            const double h2 = 2.0 * u.w;                     // 1 / rho0
            const double wx = 1.0 - u.x + (Ox - uo * u.x) * h2;
            const double wy = -u.y + (Oy - uo * u.y) * h2;
            const double wz = -u.z + (Oz - uo * u.z) * h2;
            sx = R00 * wx + R10 * wy + R20 * wz;
            sy = R01 * wx + R11 * wy + R21 * wz;
            sz = R02 * wx + R12 * wy + R22 * wz;
        };
        double a0x, a0y, a0z, a1x, a1y, a1z, max_, may_, maz_, mbx_, mby_, mbz_;
        vdir(0, a0x, a0y, a0z);
        vdir(m_last, a1x, a1y, a1z);
        vdir(m_a, max_, may_, maz_);
        vdir(m_b, mbx_, mby_, mbz_);
        const double inv_n = m_last > 0 ? 1.0 / m_last : 0.0;
        // chord through the end pixels, lowered by half its gap to the row's
        // middle (minimax line for a quadratic)
        const double gapx = 0.5 * (a0x + a1x) - 0.5 * (max_ + mbx_);
        const double gapy = 0.5 * (a0y + a1y) - 0.5 * (may_ + mby_);
        const double gapz = 0.5 * (a0z + a1z) - 0.5 * (maz_ + mbz_);
        const float v0x = static_cast<float>(a0x - 0.5 * gapx);
        const float v0y = static_cast<float>(a0y - 0.5 * gapy);
        const float v0z = static_cast<float>(a0z - 0.5 * gapz);
        const float gx = static_cast<float>((a1x - a0x) * inv_n);
        const float gy = static_cast<float>((a1y - a0y) * inv_n);
        const float gz = static_cast<float>((a1z - a0z) * inv_n);
        // 1 / (2 K rho) at the lane's middle pixel
        const float hh = SECOND ? static_cast<float>(A.u0h[pb + m_a].w / A.K) : 0.f;

        float S_r[KP], S_i[KP];
#pragma unroll
        for (int m = 0; m < KP; ++m) { S_r[m] = 0.f; S_i[m] = 0.f; }
        const AtomRec* __restrict__ at = A.atoms + static_cast<int64_t>(t) * TILE;
        const int cnt = tr.count;
        constexpr float TWO_PI_F = 6.2831853071795865f;
        AtomRec nxt = at[0];
        for (int j = 0; j < cnt; ++j) {
            const AtomRec a = nxt;                 // next record in flight meanwhile
            if (j + 1 < cnt) nxt = at[j + 1];
            float psi = fmaf(v0z, a.sz, fmaf(v0y, a.sy, v0x * a.sx));
            const float d = fmaf(gz, a.sz, fmaf(gy, a.sy, gx * a.sx));
            if (SECOND) {
                // u . sigma at the middle pixel = x_hat . sigma - v_mid . sigma
                const float u = fmaf(xs2, a.sz, fmaf(xs1, a.sy, xs0 * a.sx)) - fmaf(hmid, d, psi);
                psi = fmaf(hh, fmaf(-u, u, a.s2), psi);
            }
            const float r = __builtin_amdgcn_fractf(psi + a.ph);
            const float sn = __builtin_amdgcn_sinf(r);
            const float cs = __builtin_amdgcn_cosf(r);
            float er = a.mag * cs;
            float ei = a.mag * sn;
            const float x = d * TWO_PI_F;
            const float x2 = x * x;
            const float ar = fmaf(x2, fmaf(x2, 0.041666668f, -0.5f), 1.0f);
            const float ai = x * fmaf(x2, -0.16666667f, 1.0f);
#pragma unroll
            for (int m = 0; m < KP; ++m) {
                S_r[m] += er;
                S_i[m] += ei;
                if (m + 1 < KP) {
                    const float tr_ = fmaf(er, ar, -ei * ai);
                    ei = fmaf(er, ai, ei * ar);
                    er = tr_;
                }
            }
        }
        const float2* __restrict__ G = A.gtab + static_cast<int64_t>(tr.species) * A.npix;
#pragma unroll
        for (int m = 0; m < KP; ++m) {
            const int p = pb + min(m, m_last);
            const double4 u = A.u0h[p];
            const double uo = u.x * Ox + u.y * Oy + u.z * Oz;
            double th = A.theta0[p] + A.K * (Ox - uo + (O2 - uo * uo) * u.w);
            th -= rint(th);
            const float thf = __builtin_amdgcn_fractf(static_cast<float>(th));
            const float sn = __builtin_amdgcn_sinf(thf);
            const float cs = __builtin_amdgcn_cosf(thf);
            const float2 g = G[p];
            const float zr = g.x * cs - g.y * sn, zi = g.x * sn + g.y * cs;
            double2 c = acc[m][threadIdx.x];
            c.x += static_cast<double>(zr * S_r[m] - zi * S_i[m]);
            c.y += static_cast<double>(zr * S_i[m] + zi * S_r[m]);
            acc[m][threadIdx.x] = c;
        }
    }
    if (!lane_ok) return;
    double2* out = A.partial + (static_cast<int64_t>(frame) * A.n_slices + slice) * A.npix;
#pragma unroll
    for (int m = 0; m < KP; ++m)
        if (m < nv) out[pbase + m] = acc[m][threadIdx.x];
}

template <int KP, bool SECOND>
hipError_t launch_rec(const RecArgs& A, int n_slices, int n_frames, hipStream_t st) {
    const int lanes = A.Nz * A.nbx;
    dim3 grid(blocks_for(lanes, SUM_THREADS), n_slices, n_frames);
    hipLaunchKernelGGL((k_sum_rec<KP, SECOND>), grid, dim3(SUM_THREADS), 0, st, A);
    return hipGetLastError();
}

// ------------------------------------------------------------ probe sum
// Far-field sum for a few pixels (sparse probes of a frame): one lane per
// (frame, pixel) output, workgroups over those outputs and atom slices.  The
// row kernel gives each lane a run of KP pixels and walks the slice's atoms
// once per lane; with a few dozen pixels its grid is a handful of lanes per
// workgroup.  Here every lane of a workgroup is a different output: the
// workgroup loads each tile's atom records into LDS once (coalesced, the next
// tile's load in flight during the current one), and every lane reads them
// as broadcasts.  Per pair a lane does psi = v . sigma (v the exact direction
// at its own pixel, no line model), one sincos and the amplitude accumulate.
// Per (tile, output) the origin phase and factor are the row kernel's
// (far-field expansion, per-pixel table); the bounds are the row kernel's
// without its row line fit.
struct ProbeArgs {
    const AtomRec* atoms;
    const TileRec* tiles;
    const FrameRec* frames;
    const double4* u0h;
    const double* theta0;
    const float2* gtab;
    double2* partial;
    int n_tiles;
    int tiles_per_slice;
    int n_slices;
    int npix;
    int n_frames;
    double K;         // 1 / lambda
};

static_assert(TILE == SUM_THREADS, "one atom record per thread when a tile is loaded");
// Pixel-poses up to which the automatic choice prefers the probe form to the
// per-tile form (64 pixels x 64 poses of 200k atoms: 0.41 ms against 4.5 ms).
constexpr int64_t PROBE_AUTO_OUTPUTS = 16384;

template <bool SECOND>
__global__ __launch_bounds__(SUM_THREADS) void k_probe(const ProbeArgs A) {
    __shared__ float4 s_pos[2][TILE];     // sx, sy, sz, |s|^2
    __shared__ float2 s_amp[2][TILE];     // |w|, arg(w) / 2 pi
    const int slice = blockIdx.y;
    const int t0 = slice * A.tiles_per_slice;
    const int t1 = min(A.n_tiles, t0 + A.tiles_per_slice);
    const int64_t o = blockIdx.x * static_cast<int64_t>(SUM_THREADS) + threadIdx.x;
    const bool ok = o < static_cast<int64_t>(A.npix) * A.n_frames;
    const int frame = ok ? static_cast<int>(o / A.npix) : 0;
    const int p = ok ? static_cast<int>(o % A.npix) : 0;
    const FrameRec& F = A.frames[frame];
    const double R00 = F.R[0], R01 = F.R[1], R02 = F.R[2];
    const double R10 = F.R[3], R11 = F.R[4], R12 = F.R[5];
    const double R20 = F.R[6], R21 = F.R[7], R22 = F.R[8];
    const float xs0 = static_cast<float>(F.xs[0]);
    const float xs1 = static_cast<float>(F.xs[1]);
    const float xs2 = static_cast<float>(F.xs[2]);
    const double4 u = A.u0h[p];
    const double th0 = A.theta0[p];
    const double h2 = 2.0 * u.w;                                  // 1 / rho0
    const float hh = SECOND ? static_cast<float>(u.w / A.K) : 0.f;  // 1 / (2 K rho0)
    double acc_r = 0.0, acc_i = 0.0;
    if (t0 >= t1) {
        if (ok) A.partial[(static_cast<int64_t>(frame) * A.n_slices + slice) * A.npix + p] =
                    make_double2(0.0, 0.0);
        return;
    }
    // records past a tile's count stay zero (|w| = 0), so every lane can
    // walk the full tile without a count-dependent loop bound
    auto fetch = [&](int t, float4& ps, float2& am) {
        const AtomRec r = A.atoms[static_cast<int64_t>(t) * TILE + threadIdx.x];
        ps = make_float4(r.sx, r.sy, r.sz, r.s2);
        am = make_float2(r.mag, r.ph);
    };
    float4 nps;
    float2 nam;
    fetch(t0, nps, nam);
    s_pos[0][threadIdx.x] = nps;
    s_amp[0][threadIdx.x] = nam;
    __syncthreads();
    for (int t = t0; t < t1; ++t) {
        const int cur = (t - t0) & 1;
        if (t + 1 < t1) fetch(t + 1, nps, nam);          // in flight through this tile
        const TileRec tr = A.tiles[t];
        const double Ox = fma(R00, tr.ox, fma(R01, tr.oy, fma(R02, tr.oz, F.T[0])));
        const double Oy = fma(R10, tr.ox, fma(R11, tr.oy, fma(R12, tr.oz, F.T[1])));
        const double Oz = fma(R20, tr.ox, fma(R21, tr.oy, fma(R22, tr.oz, F.T[2])));
        const double O2 = Ox * Ox + Oy * Oy + Oz * Oz;
        const double uo = u.x * Ox + u.y * Oy + u.z * Oz;
        // x_hat minus the direction from the tile origin to the pixel, to
        // first order in |O| / rho0, in the sample frame
        const double wx = 1.0 - u.x + (Ox - uo * u.x) * h2;
        const double wy = -u.y + (Oy - uo * u.y) * h2;
        const double wz = -u.z + (Oz - uo * u.z) * h2;
        const float vx = static_cast<float>(R00 * wx + R10 * wy + R20 * wz);
        const float vy = static_cast<float>(R01 * wx + R11 * wy + R21 * wz);
        const float vz = static_cast<float>(R02 * wx + R12 * wy + R22 * wz);
        float S_r = 0.f, S_i = 0.f;
#pragma unroll 8
        for (int j = 0; j < TILE; ++j) {
            const float4 a = s_pos[cur][j];
            const float2 w = s_amp[cur][j];
            float psi = fmaf(vz, a.z, fmaf(vy, a.y, vx * a.x));
            if (SECOND) {
                const float us = fmaf(xs2, a.z, fmaf(xs1, a.y, xs0 * a.x)) - psi;
                psi = fmaf(hh, fmaf(-us, us, a.w), psi);
            }
            const float r = __builtin_amdgcn_fractf(psi + w.y);
            S_r = fmaf(w.x, __builtin_amdgcn_cosf(r), S_r);
            S_i = fmaf(w.x, __builtin_amdgcn_sinf(r), S_i);
        }
        double th = th0 + A.K * (Ox - uo + (O2 - uo * uo) * u.w);
        th -= rint(th);
        const float thf = __builtin_amdgcn_fractf(static_cast<float>(th));
        const float sn = __builtin_amdgcn_sinf(thf), cs = __builtin_amdgcn_cosf(thf);
        const float2 g = A.gtab[static_cast<int64_t>(tr.species) * A.npix + p];
        const float zr = g.x * cs - g.y * sn, zi = g.x * sn + g.y * cs;
        acc_r += static_cast<double>(zr * S_r - zi * S_i);
        acc_i += static_cast<double>(zr * S_i + zi * S_r);
        if (t + 1 < t1) {
            s_pos[cur ^ 1][threadIdx.x] = nps;           // the other buffer: last read a
            s_amp[cur ^ 1][threadIdx.x] = nam;           // tile ago, before the barrier
        }
        __syncthreads();
    }
    if (ok)
        A.partial[(static_cast<int64_t>(frame) * A.n_slices + slice) * A.npix + p] =
            make_double2(acc_r, acc_i);
}

template <bool SECOND>
hipError_t launch_probe(const ProbeArgs& A, int n_slices, hipStream_t st) {
    const int64_t outs = static_cast<int64_t>(A.npix) * A.n_frames;
    dim3 grid(blocks_for(outs, SUM_THREADS), n_slices);
    hipLaunchKernelGGL((k_probe<SECOND>), grid, dim3(SUM_THREADS), 0, st, A);
    return hipGetLastError();
}

// ------------------------------------------------------- matrix-core sum
// Far-field sum on the matrix cores.  One wave owns a patch of MMA_P x MMA_P
// pixels (a along the row, b down the columns, both offsets from the patch's
// middle pixel).  Across the patch the atom-offset direction is modelled per
// axis to second order,
//     v(a, b) = va(a) + wb(b),  va(a) = vc + a gy + a^2 hy,  wb(b) = b gz + b^2 hz,
// exact along the middle row and column (the host bounds the deviation
// elsewhere from the measured pixel grid), so per atom the phasor separates,
//     e^{2 pi i v.sigma} = A(a) B(b),  A(a) = w e^{2 pi i va(a).sigma},
//     B(b) = e^{2 pi i wb(b).sigma},
// and the tile sum S(a, b) = sum_j A_j(a) B_j(b) is a complex matrix product
// over the atoms.  With A = Ar + i Ai and B = Br + i Bi, four real products
// give it, S = (Ar.Br - Ai.Bi) + i (Ar.Bi + Ai.Br), accumulated register by
// register in the same lane.  The instruction is v_mfma_f32_32x32x16_f16 (16
// atoms per step; lane l holds row / column l mod 32 and atoms
// 8 (l / 32) .. + 7, staged in LDS per tile), and each operand value is split
// into two FP16 halves (22 significant bits), so three products per real
// product keep FP32 accuracy: hi.hi + hi.lo + lo.hi, 12 MFMAs per step.
// Each lane forms, per atom, the A value of its row and the B value of its
// column (one v_sin and one v_cos each); |w| enters as sqrt|w| on both, so
// every operand is a product that the FP16 split handles in one mixed FMA.
//
// The per-(tile, pixel) origin phase is Theta0(p) + K (X_t - u_p.O +
// (|O|^2 - (u_p.O)^2) / (2 rho)).  Its pixel-only part, Theta0(p) -
// Theta0(middle), sits in the table gmma with the factor G; the tile part at
// the middle pixel is one FP64 value per tile; what remains,
// -K (u_p - u_mid).O (< 1 rev), is an FP32 dot product per pixel and tile.
// Tile results are summed in FP32 over MMA_FLUSH tiles, then added to the
// FP64 slice partial (owned by this wave).
struct MmaArgs {
    const AtomRec* atoms;
    const TileRec* tiles;
    const FrameRec* frames;
    const double4* u0h;
    const double* theta0;
    const float4* dup;
    const float2* gmma;
    double2* partial;
    int n_tiles;
    int tiles_per_slice;
    int n_slices;
    int Ny;
    int Nz;
    int npix;
    int pby;          // patches along a row
    int n_patches;
    double K;         // 1 / lambda
    const float* grp_r;   // moment sum: radii of the pairs / quads of tiles ([tile][2])
    float grp_rmax;       // largest unit radius inside the error bounds (0: tiles alone)
};

constexpr int MMA_FLUSH = 32;
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

// x = hi + lo, both FP16: 22 significant bits
__device__ __forceinline__ void split16(float x, halfx8& hi, halfx8& lo, int u) {
    const _Float16 h = static_cast<_Float16>(x);
    hi[u] = h;
    lo[u] = static_cast<_Float16>(x - static_cast<float>(h));
}

template <bool SECOND>
__global__ __launch_bounds__(64 * MMA_WAVES) void k_sum_mma(const MmaArgs A) {
    // the workgroup's current and next tile (all its waves walk the same
    // tiles), and the FP32 running sums of each lane's 16 pixels
    __shared__ AtomRec tile_buf[2][TILE];
    __shared__ float2 lacc[MMA_WAVES][16][64];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int patch = blockIdx.x * MMA_WAVES + wave;
    const bool live = patch < A.n_patches;       // idle waves still take part in the barriers
    const int frame = blockIdx.z;
    const int slice = blockIdx.y;
    const FrameRec& F = A.frames[frame];
    const int t0 = slice * A.tiles_per_slice;
    const int t1 = min(A.n_tiles, t0 + A.tiles_per_slice);
    const int pc = live ? patch : 0;
    const PatchGeom P = patch_of(pc % A.pby, pc / A.pby, A.Ny, A.Nz);
    const int Ny = A.Ny;

    // operand roles: A row a = q and B column b = q; this lane's atoms in a
    // 16-atom step are 8 g .. 8 g + 7
    const int g = lane >> 5;
    const int q = lane & 31;
    const double R00 = F.R[0], R01 = F.R[1], R02 = F.R[2];
    const double R10 = F.R[3], R11 = F.R[4], R12 = F.R[5];
    const double R20 = F.R[6], R21 = F.R[7], R22 = F.R[8];
    const int pmid = (P.z0 + P.mz) * Ny + P.y0 + P.my;

    // direction model for the lab origin, exact at the middle pixel and the
    // ends of the middle row and column (sample frame)
    double va0[3], vc0[3];
    float wb[3];
    {
        const int samp[5] = {pmid, pmid - P.my, pmid - P.my + P.vy - 1, pmid - P.mz * Ny,
                             pmid + (P.vz - 1 - P.mz) * Ny};
        double vs[5][3];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const double4 u = A.u0h[samp[k]];
            const double wx = 1.0 - u.x, wy = -u.y, wz = -u.z;
            vs[k][0] = R00 * wx + R10 * wy + R20 * wz;
            vs[k][1] = R01 * wx + R11 * wy + R21 * wz;
            vs[k][2] = R02 * wx + R12 * wy + R22 * wz;
        }
        const double fa = q - P.my, fb = q - P.mz;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double gq, hq;
            quad_fit(vs[1][d] - vs[0][d], vs[2][d] - vs[0][d], -P.my, P.vy - 1 - P.my, gq, hq);
            va0[d] = vs[0][d] + fa * (gq + fa * hq);
            vc0[d] = vs[0][d];
            quad_fit(vs[3][d] - vs[0][d], vs[4][d] - vs[0][d], -P.mz, P.vz - 1 - P.mz, gq, hq);
            wb[d] = static_cast<float>(fb * (gq + fb * hq));
        }
    }
    const double4 um = A.u0h[pmid];              // middle pixel: direction, 1 / (2 rho)
    const double th_mid = A.theta0[pmid];
    // u_hat . sigma at the middle pixel = (x_hat - v_mid) . sigma (second order)
    const float xc0 = static_cast<float>(F.xs[0] - vc0[0]);
    const float xc1 = static_cast<float>(F.xs[1] - vc0[1]);
    const float xc2 = static_cast<float>(F.xs[2] - vc0[2]);
    const float hh = SECOND ? static_cast<float>(um.w / A.K) : 0.f;

    // result register r of this lane: pixel row a(r) = 8 (r / 4) + 4 g + r % 4,
    // column q
    const int prow = (P.z0 + min(q, P.vz - 1)) * Ny + P.y0;
    auto arow = [&](int r) { return 8 * (r >> 2) + 4 * g + (r & 3); };
    float2 (*acc)[64] = lacc[wave];
#pragma unroll
    for (int r = 0; r < 16; ++r) acc[r][lane] = make_float2(0.f, 0.f);
    double2* out = A.partial + (static_cast<int64_t>(frame) * A.n_slices + slice) * A.npix;
    bool first = true;
    int since = 0;

    if (t0 < t1) tile_buf[0][tid] = A.atoms[static_cast<int64_t>(t0) * TILE + tid];
    __syncthreads();
    for (int t = t0; t < t1; ++t) {
        const int buf = (t - t0) & 1;
        const bool has_next = t + 1 < t1;
        // this thread's record of the next tile, in flight during this one
        const float4* nsrc = reinterpret_cast<const float4*>(
            A.atoms + static_cast<int64_t>(has_next ? t + 1 : t) * TILE + tid);
        const float4 n0 = nsrc[0], n1 = nsrc[1];
        const TileRec tr = A.tiles[t];
        if (live) {
            // keeps the table loads below inside the loop (count <= 256: shift 0)
            const int tie = tr.count >> 16;
            const double Ox = fma(R00, tr.ox, fma(R01, tr.oy, fma(R02, tr.oz, F.T[0])));
            const double Oy = fma(R10, tr.ox, fma(R11, tr.oy, fma(R12, tr.oz, F.T[1])));
            const double Oz = fma(R20, tr.ox, fma(R21, tr.oy, fma(R22, tr.oz, F.T[2])));
            const double O2 = Ox * Ox + Oy * Oy + Oz * Oz;
            const double uo = um.x * Ox + um.y * Oy + um.z * Oz;
            double ph_t = th_mid + A.K * (Ox - uo + (O2 - uo * uo) * um.w);
            ph_t -= floor(ph_t);
            // first-order change of the direction for this tile origin,
            // constant over the patch: R^T (O - (u.O) u) / rho
            const double h2 = 2.0 * um.w;
            const double cxl = (Ox - uo * um.x) * h2, cyl = (Oy - uo * um.y) * h2,
                         czl = (Oz - uo * um.z) * h2;
            const float va_x = static_cast<float>(va0[0] + (R00 * cxl + R10 * cyl + R20 * czl));
            const float va_y = static_cast<float>(va0[1] + (R01 * cxl + R11 * cyl + R21 * czl));
            const float va_z = static_cast<float>(va0[2] + (R02 * cxl + R12 * cyl + R22 * czl));
            const float kox = static_cast<float>(A.K * Ox), koy = static_cast<float>(A.K * Oy),
                        koz = static_cast<float>(A.K * Oz);
            const float pht = static_cast<float>(ph_t);

            floatx16 crr, cii, cri, cir;
#pragma unroll
            for (int r = 0; r < 16; ++r) { crr[r] = 0.f; cii[r] = 0.f; cri[r] = 0.f; cir[r] = 0.f; }
            const AtomRec* at = tile_buf[buf] + 8 * g;
            const int cnt = tr.count;
            // records past the count are zero (mag = 0); j + 15 <= 255
#pragma unroll 1
            for (int j = 0; j < cnt; j += 16) {
                halfx8 arh, arl, aih, ail, brh, brl, bih, bil;
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const AtomRec a = at[j + u];
                    float pa = fmaf(va_z, a.sz, fmaf(va_y, a.sy, va_x * a.sx));
                    if (SECOND) {
                        const float w = fmaf(xc2, a.sz, fmaf(xc1, a.sy, xc0 * a.sx));
                        pa = fmaf(hh, fmaf(-w, w, a.s2), pa);
                    }
                    // phases in revolutions; v_sin / v_cos reduce |x| <= 256
                    // themselves (the host bounds |pa| + 1 by the tile radius)
                    const float fa = pa + a.ph;
                    const float fb = fmaf(wb[2], a.sz, fmaf(wb[1], a.sy, wb[0] * a.sx));
                    // |w| split as sqrt|w| on each operand: every value is a
                    // product, which the FP16 split turns into one mixed FMA
                    split16(a.rmag * __builtin_amdgcn_cosf(fa), arh, arl, u);
                    split16(a.rmag * __builtin_amdgcn_sinf(fa), aih, ail, u);
                    split16(a.rmag * __builtin_amdgcn_cosf(fb), brh, brl, u);
                    split16(a.rmag * __builtin_amdgcn_sinf(fb), bih, bil, u);
                }
                crr = __builtin_amdgcn_mfma_f32_32x32x16_f16(arh, brh, crr, 0, 0, 0);
                cii = __builtin_amdgcn_mfma_f32_32x32x16_f16(aih, bih, cii, 0, 0, 0);
                cri = __builtin_amdgcn_mfma_f32_32x32x16_f16(arh, bih, cri, 0, 0, 0);
                cir = __builtin_amdgcn_mfma_f32_32x32x16_f16(aih, brh, cir, 0, 0, 0);
                crr = __builtin_amdgcn_mfma_f32_32x32x16_f16(arh, brl, crr, 0, 0, 0);
                cii = __builtin_amdgcn_mfma_f32_32x32x16_f16(aih, bil, cii, 0, 0, 0);
                cri = __builtin_amdgcn_mfma_f32_32x32x16_f16(arh, bil, cri, 0, 0, 0);
                cir = __builtin_amdgcn_mfma_f32_32x32x16_f16(aih, brl, cir, 0, 0, 0);
                crr = __builtin_amdgcn_mfma_f32_32x32x16_f16(arl, brh, crr, 0, 0, 0);
                cii = __builtin_amdgcn_mfma_f32_32x32x16_f16(ail, bih, cii, 0, 0, 0);
                cri = __builtin_amdgcn_mfma_f32_32x32x16_f16(arl, bih, cri, 0, 0, 0);
                cir = __builtin_amdgcn_mfma_f32_32x32x16_f16(ail, brh, cir, 0, 0, 0);
            }

            // S(a(r), q) = (C_rr - C_ii) + i (C_ri + C_ir); times the origin
            // phasor and G of the pixel
            const float2* __restrict__ Gm = A.gmma + static_cast<int64_t>(tr.species) * A.npix;
            float s_re[16], s_im[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                s_re[r] = crr[r] - cii[r];
                s_im[r] = cri[r] + cir[r];
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                // four pixels at a time: bounds the table loads in flight
                if ((r & 3) == 0) __builtin_amdgcn_sched_barrier(0);
                const float sr = s_re[r], si = s_im[r];
                const int p = prow + min(arow(r), P.vy - 1) + tie;
                const float4 du = A.dup[p];
                const float phs = __builtin_amdgcn_fractf(
                    pht - fmaf(koz, du.z, fmaf(koy, du.y, kox * du.x)));
                const float sn = __builtin_amdgcn_sinf(phs);
                const float cs = __builtin_amdgcn_cosf(phs);
                const float2 gg = Gm[p];
                const float zr = gg.x * cs - gg.y * sn, zi = gg.x * sn + gg.y * cs;
                float2 av = acc[r][lane];
                av.x = fmaf(zr, sr, fmaf(-zi, si, av.x));
                av.y = fmaf(zr, si, fmaf(zi, sr, av.y));
                acc[r][lane] = av;
            }
            if (++since == MMA_FLUSH || t + 1 == t1) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    if ((r & 3) == 0) __builtin_amdgcn_sched_barrier(0);
                    const float2 av = acc[r][lane];
                    if (arow(r) < P.vy && q < P.vz) {
                        double2* o = out + prow + arow(r);
                        double2 v = first ? make_double2(0.0, 0.0) : *o;
                        v.x += static_cast<double>(av.x);
                        v.y += static_cast<double>(av.y);
                        *o = v;
                    }
                    acc[r][lane] = make_float2(0.f, 0.f);
                }
                first = false;
                since = 0;
            }
        }
        // the other buffer was last read in the previous tile, which every
        // wave finished before the previous barrier
        if (has_next) {
            float4* dst = reinterpret_cast<float4*>(&tile_buf[buf ^ 1][tid]);
            dst[0] = n0;
            dst[1] = n1;
        }
        __syncthreads();
    }
    if (live && first) {                          // a slice without tiles
#pragma unroll
        for (int r = 0; r < 16; ++r)
            if (arow(r) < P.vy && q < P.vz) out[prow + arow(r)] = make_double2(0.0, 0.0);
    }
}

template <bool SECOND>
hipError_t launch_mma(const MmaArgs& A, int n_slices, int n_frames, hipStream_t st) {
    dim3 grid((A.n_patches + MMA_WAVES - 1) / MMA_WAVES, n_slices, n_frames);
    hipLaunchKernelGGL((k_sum_mma<SECOND>), grid, dim3(64 * MMA_WAVES), 0, st, A);
    return hipGetLastError();
}

// ------------------------------------------------------------- moment sum
// Far-field sum by moments over the same 32 x 32 patches and direction model
// as the matrix-core sum, with the model's cross term added.  With x = a / 16
// and y = b / 16 (|x|, |y| <= 1) the phase of atom j at pixel (a, b), relative
// to the patch's middle pixel, is
//     2 pi theta_j + P_j x + Q_j x^2 + R_j y + S_j y^2 + U_j x y   [rad],
//     theta_j = va(0).sigma_j + ph_j,  P_j = 2 pi 16 gy.sigma_j,  Q_j = 2 pi 256 hy.sigma_j,
//     R_j = 2 pi 16 gz.sigma_j,  S_j = 2 pi 256 hz.sigma_j,  U_j = 2 pi 256 gyz.sigma_j,
// gyz the mixed difference of the patch's corner directions (the separable
// model leaves a b gyz out; with it the model's deviation drops about tenfold
// on a flat detector).  The host bounds phi = max |P x + R y| (a few tenths of
// a radian: the radius times the patch's angular size), |Q|, |S| (~1e-4 rad)
// and |U| (~1e-5 rad), so
//     e^{i(P x + R y)} (1 + i Q x^2 + i S y^2 + i U x y)
//         = sum_{m+n<=D} i^{m+n} [T_mn - i (Q T_{m-2,n} + S T_{m,n-2} + U T_{m-1,n-1})] x^m y^n,
//     T_mn = P^m R^n / (m! n!),
// to phi^{D+1} / (D+1)! (the Q, S terms kept where the rest of their monomial
// has degree <= 3, the U terms where it has degree <= 1).  The sum over a unit
// of 1, 2 or 4 tiles (their atoms' offsets taken from the unit's origin) over
// the patch is then a polynomial,
//     S_t(x, y) = sum_{m+n<=D} C_mn x^m y^n,  C_mn = sum_j |w_j| e^{2 pi i theta_j} (...)_mn,
// whose (D+1)(D+2)/2 complex moments cost one sincos and ~50 FMAs per atom
// and patch (the matrix-core form builds 64 phasors per atom and patch).
// Each lane accumulates the moments of 4 atoms of each tile of the unit; at
// the unit's last tile the wave sums them through LDS; every lane then
// evaluates its 16 pixels by Horner's rule and applies the per-(unit, pixel)
// origin phase as the matrix-core sum does per tile; the pixel's table factor (a function of species and pixel only)
// multiplies the running sums when they are flushed.
constexpr int POLY_WAVES = 4;
// Units summed in FP32 between flushes into the FP64 slice partial (each
// flush reads and writes 16 B per pixel of the patch).
constexpr int POLY_FLUSH = 128;


template <int D>
struct PolyN {
    static constexpr int NC = (D + 1) * (D + 2) / 2;   // moments m + n <= D
    static constexpr int NV = 2 * NC;                   // floats
    static constexpr int NVP = NV | 1;                  // odd row stride in LDS: one
                                                        // bank per lane for the stores
};

template <int D, bool SECOND>
__global__ __launch_bounds__(64 * POLY_WAVES) __attribute__((amdgpu_waves_per_eu(2)))
void k_sum_poly(const MmaArgs A) {
    constexpr int NC = PolyN<D>::NC, NVP = PolyN<D>::NVP;
    constexpr float TWO_PI_F = 6.283185307179586f;
    __shared__ AtomRec tile_buf[2][TILE];
    __shared__ float red[POLY_WAVES][64 * NVP];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int patch = blockIdx.x * POLY_WAVES + wave;
    const bool live = patch < A.n_patches;       // idle waves compute on patch 0, write nothing
    const int frame = blockIdx.z;
    const int slice = blockIdx.y;
    const FrameRec& F = A.frames[frame];
    const int t0 = slice * A.tiles_per_slice;
    const int t1 = min(A.n_tiles, t0 + A.tiles_per_slice);
    const int pc = live ? patch : 0;
    const PatchGeom P = patch_of(pc % A.pby, pc / A.pby, A.Ny, A.Nz);
    const int Ny = A.Ny;
    const int g = lane >> 5;
    const int q = lane & 31;
    const double R00 = F.R[0], R01 = F.R[1], R02 = F.R[2];
    const double R10 = F.R[3], R11 = F.R[4], R12 = F.R[5];
    const double R20 = F.R[6], R21 = F.R[7], R22 = F.R[8];
    const int pmid = (P.z0 + P.mz) * Ny + P.y0 + P.my;

    // the patch's direction model (sample frame), as in the matrix-core sum
    double vc0[3];
    float gyk[3], hyk[3], gzk[3], hzk[3], uyz[3];
    {
        const int pc = pmid - P.my - P.mz * Ny;
        const int samp[9] = {pmid, pmid - P.my, pmid - P.my + P.vy - 1, pmid - P.mz * Ny,
                             pmid + (P.vz - 1 - P.mz) * Ny, pc, pc + P.vy - 1,
                             pc + (P.vz - 1) * Ny, pc + P.vy - 1 + (P.vz - 1) * Ny};
        double vs[9][3];
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const double4 u = A.u0h[samp[k]];
            const double wx = 1.0 - u.x, wy = -u.y, wz = -u.z;
            vs[k][0] = R00 * wx + R10 * wy + R20 * wz;
            vs[k][1] = R01 * wx + R11 * wy + R21 * wz;
            vs[k][2] = R02 * wx + R12 * wy + R22 * wz;
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double gq, hq;
            vc0[d] = vs[0][d];
            quad_fit(vs[1][d] - vs[0][d], vs[2][d] - vs[0][d], -P.my, P.vy - 1 - P.my, gq, hq);
            gyk[d] = static_cast<float>(2.0 * PI * 16.0 * gq);
            hyk[d] = static_cast<float>(2.0 * PI * 256.0 * hq);
            quad_fit(vs[3][d] - vs[0][d], vs[4][d] - vs[0][d], -P.mz, P.vz - 1 - P.mz, gq, hq);
            gzk[d] = static_cast<float>(2.0 * PI * 16.0 * gq);
            hzk[d] = static_cast<float>(2.0 * PI * 256.0 * hq);
            // cross term a b g_yz from the mixed difference of the corners
            const double area = static_cast<double>(P.vy - 1) * static_cast<double>(P.vz - 1);
            const double gyz = area > 0.0 ? (vs[8][d] - vs[7][d] - vs[6][d] + vs[5][d]) / area : 0.0;
            uyz[d] = static_cast<float>(2.0 * PI * 256.0 * gyz);
        }
    }
    const double4 um = A.u0h[pmid];
    const double th_mid = A.theta0[pmid];
    const float xc0 = static_cast<float>(F.xs[0] - vc0[0]);
    const float xc1 = static_cast<float>(F.xs[1] - vc0[1]);
    const float xc2 = static_cast<float>(F.xs[2] - vc0[2]);
    const float hh = SECOND ? static_cast<float>(um.w / A.K) : 0.f;

    // this lane's pixels: rows a(r) = 8 (r / 4) + 4 g + r % 4 of column q
    const int prow = (P.z0 + min(q, P.vz - 1)) * Ny + P.y0;
    auto arow = [&](int r) { return 8 * (r >> 2) + 4 * g + (r & 3); };
    const float yq = static_cast<float>(q - P.mz) * (1.f / 16.f);
    float acc_re[16], acc_im[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) { acc_re[r] = 0.f; acc_im[r] = 0.f; }
    double2* out = A.partial + (static_cast<int64_t>(frame) * A.n_slices + slice) * A.npix;
    bool first = true;
    int since = 0;
    int acc_species = 0;
    float* rw = red[wave];
    // running sums times the pixel table of their (single) species, into the
    // FP64 slice partial
    auto flush = [&]() {
        const float2* __restrict__ Gm = A.gmma + static_cast<int64_t>(acc_species) * A.npix;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            if (live && arow(r) < P.vy && q < P.vz) {
                const int p = prow + arow(r);
                const float2 gg = Gm[p];
                double2* o = out + p;
                double2 v = first ? make_double2(0.0, 0.0) : *o;
                v.x += static_cast<double>(gg.x * acc_re[r] - gg.y * acc_im[r]);
                v.y += static_cast<double>(gg.x * acc_im[r] + gg.y * acc_re[r]);
                *o = v;
            }
            acc_re[r] = 0.f;
            acc_im[r] = 0.f;
        }
        first = false;
        since = 0;
    };

    // the running unit: 4, 2 or 1 consecutive tiles of one species expanded
    // about one origin (the staging measured the units' radii)
    float cr[NC], ci[NC];
    int left = 0, tie = 0;
    double uox = 0.0, uoy = 0.0, uoz = 0.0;
    float vx = 0.f, vy = 0.f, vz = 0.f, kox = 0.f, koy = 0.f, koz = 0.f, pht = 0.f;
    if (t0 < t1) tile_buf[0][tid] = A.atoms[static_cast<int64_t>(t0) * TILE + tid];
    __syncthreads();
    for (int t = t0; t < t1; ++t) {
        const int buf = (t - t0) & 1;
        const bool has_next = t + 1 < t1;
        const float4* nsrc = reinterpret_cast<const float4*>(
            A.atoms + static_cast<int64_t>(has_next ? t + 1 : t) * TILE + tid);
        const float4 n0 = nsrc[0], n1 = nsrc[1];
        const TileRec tr = A.tiles[t];
        if (left == 0) {
            int gsz = 1;
            if (A.grp_rmax > 0.f) {
                if (t + 3 < t1 && A.grp_r[2 * static_cast<int64_t>(t) + 1] <= A.grp_rmax) gsz = 4;
                else if (t + 1 < t1 && A.grp_r[2 * static_cast<int64_t>(t)] <= A.grp_rmax) gsz = 2;
            }
            tie = tr.count >> 16;
            // tiles are grouped by species; the sums carry one species' table
            if (since > 0 && tr.species != acc_species) flush();
            acc_species = tr.species;
            uox = tr.ox;
            uoy = tr.oy;
            uoz = tr.oz;
            if (gsz > 1) {
                const TileRec b1 = A.tiles[t + 1];
                const TileRec b2 = A.tiles[gsz == 4 ? t + 2 : t];
                const TileRec b3 = A.tiles[gsz == 4 ? t + 3 : t];
                uox = unit_origin(gsz, tr.ox, b1.ox, b2.ox, b3.ox);
                uoy = unit_origin(gsz, tr.oy, b1.oy, b2.oy, b3.oy);
                uoz = unit_origin(gsz, tr.oz, b1.oz, b2.oz, b3.oz);
            }
            const double Ox = fma(R00, uox, fma(R01, uoy, fma(R02, uoz, F.T[0])));
            const double Oy = fma(R10, uox, fma(R11, uoy, fma(R12, uoz, F.T[1])));
            const double Oz = fma(R20, uox, fma(R21, uoy, fma(R22, uoz, F.T[2])));
            const double O2 = Ox * Ox + Oy * Oy + Oz * Oz;
            const double uo = um.x * Ox + um.y * Oy + um.z * Oz;
            double ph_t = th_mid + A.K * (Ox - uo + (O2 - uo * uo) * um.w);
            ph_t -= floor(ph_t);
            const double h2 = 2.0 * um.w;
            const double cxl = (Ox - uo * um.x) * h2, cyl = (Oy - uo * um.y) * h2,
                         czl = (Oz - uo * um.z) * h2;
            vx = static_cast<float>(vc0[0] + (R00 * cxl + R10 * cyl + R20 * czl));
            vy = static_cast<float>(vc0[1] + (R01 * cxl + R11 * cyl + R21 * czl));
            vz = static_cast<float>(vc0[2] + (R02 * cxl + R12 * cyl + R22 * czl));
            kox = static_cast<float>(A.K * Ox);
            koy = static_cast<float>(A.K * Oy);
            koz = static_cast<float>(A.K * Oz);
            pht = static_cast<float>(ph_t);
#pragma unroll
            for (int k = 0; k < NC; ++k) { cr[k] = 0.f; ci[k] = 0.f; }
            left = gsz;
        }
        // this tile's offset from the unit origin [wavelengths, sample frame]
        const float dx = static_cast<float>((tr.ox - uox) * A.K);
        const float dy = static_cast<float>((tr.oy - uoy) * A.K);
        const float dz = static_cast<float>((tr.oz - uoz) * A.K);

        // moments of this lane's 4 atoms (records past the count are zero)
        const AtomRec* at = tile_buf[buf];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const AtomRec a = at[lane + 64 * i];
            const float sx = a.sx + dx, sy = a.sy + dy, sz = a.sz + dz;
            float th = fmaf(vz, sz, fmaf(vy, sy, vx * sx)) + a.ph;
            if (SECOND) {
                const float w = fmaf(xc2, sz, fmaf(xc1, sy, xc0 * sx));
                const float s2 = fmaf(sz, sz, fmaf(sy, sy, sx * sx));
                th = fmaf(hh, fmaf(-w, w, s2), th);
            }
            const float wr = a.mag * __builtin_amdgcn_cosf(th);
            const float wi = a.mag * __builtin_amdgcn_sinf(th);
            const float Pj = fmaf(gyk[2], sz, fmaf(gyk[1], sy, gyk[0] * sx));
            const float Rj = fmaf(gzk[2], sz, fmaf(gzk[1], sy, gzk[0] * sx));
            const float Qj = fmaf(hyk[2], sz, fmaf(hyk[1], sy, hyk[0] * sx));
            const float Sj = fmaf(hzk[2], sz, fmaf(hzk[1], sy, hzk[0] * sx));
            const float Uj = fmaf(uyz[2], sz, fmaf(uyz[1], sy, uyz[0] * sx));
            // w P^m / m! (ar, ai) against R^n / n! (bn): two FMAs per moment;
            // the corrections -i w (Q T_{m-2,n} + S T_{m,n-2} + U T_{m-1,n-1})
            // against Q R^n / n!, S R^n / n!, U R^n / n!
            float bn[D + 1], ar[D + 1], ai[D + 1];
            bn[0] = 1.f;
            ar[0] = wr;
            ai[0] = wi;
#pragma unroll
            for (int m = 1; m <= D; ++m) {
                const float pq = Pj * (1.f / m);
                ar[m] = ar[m - 1] * pq;
                ai[m] = ai[m - 1] * pq;
                bn[m] = bn[m - 1] * Rj * (1.f / m);
            }
            float qb[4], sb[4], ub[2];
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                qb[n] = Qj * bn[n];
                sb[n] = Sj * bn[n];
                if (n < 2) ub[n] = Uj * bn[n];
            }
            int k = 0;
#pragma unroll
            for (int m = 0; m <= D; ++m) {
#pragma unroll
                for (int n = 0; n <= D - m; ++n, ++k) {
                    float xr = fmaf(ar[m], bn[n], cr[k]);
                    float xi = fmaf(ai[m], bn[n], ci[k]);
                    if (m >= 2 && m - 2 + n <= 3) {
                        xr = fmaf(ai[m - 2], qb[n], xr);
                        xi = fmaf(-ar[m - 2], qb[n], xi);
                    }
                    if (n >= 2 && m + n - 2 <= 3) {
                        xr = fmaf(ai[m], sb[n - 2], xr);
                        xi = fmaf(-ar[m], sb[n - 2], xi);
                    }
                    if (m >= 1 && n >= 1 && m + n - 2 <= 1) {
                        xr = fmaf(ai[m - 1], ub[n - 1], xr);
                        xi = fmaf(-ar[m - 1], ub[n - 1], xi);
                    }
                    cr[k] = xr;
                    ci[k] = xi;
                }
            }
        }
        if (--left > 0) {                          // the unit goes on in the next tile
            if (has_next) {
                float4* dst = reinterpret_cast<float4*>(&tile_buf[buf ^ 1][tid]);
                dst[0] = n0;
                dst[1] = n1;
            }
            __syncthreads();
            continue;
        }
        // sum over the wave's lanes: row per lane in LDS, column sums per lane
#pragma unroll
        for (int k = 0; k < NC; ++k) {
            rw[lane * NVP + 2 * k] = cr[k];
            rw[lane * NVP + 2 * k + 1] = ci[k];
        }
        __syncthreads();
        const int col = min(lane, NVP - 1);
        float tot = 0.f;
#pragma unroll 8
        for (int l = 0; l < 64; ++l) tot += rw[l * NVP + col];
        // (rw is rewritten only after the barrier that ends this tile)
        // every lane gets the moments (wave-uniform), with the i^{m+n} factor
        float Cr[NC], Ci[NC];
        {
            int k = 0;
#pragma unroll
            for (int m = 0; m <= D; ++m) {
#pragma unroll
                for (int n = 0; n <= D - m; ++n, ++k) {
                    const float re = __int_as_float(
                        __builtin_amdgcn_readlane(__float_as_int(tot), 2 * k));
                    const float im = __int_as_float(
                        __builtin_amdgcn_readlane(__float_as_int(tot), 2 * k + 1));
                    switch ((m + n) & 3) {
                        case 0: Cr[k] = re; Ci[k] = im; break;
                        case 1: Cr[k] = -im; Ci[k] = re; break;
                        case 2: Cr[k] = -re; Ci[k] = -im; break;
                        default: Cr[k] = im; Ci[k] = -re; break;
                    }
                }
            }
        }
        // G_m(y) = sum_n C_mn y^n for this lane's column, then each pixel row
        float Gr[D + 1], Gi[D + 1];
        {
            int base = 0;
#pragma unroll
            for (int m = 0; m <= D; ++m) {
                float gr = 0.f, gi = 0.f;
#pragma unroll
                for (int n = D - m; n >= 0; --n) {
                    gr = fmaf(gr, yq, Cr[base + n]);
                    gi = fmaf(gi, yq, Ci[base + n]);
                }
                Gr[m] = gr;
                Gi[m] = gi;
                base += D - m + 1;
            }
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            if ((r & 3) == 0) __builtin_amdgcn_sched_barrier(0);
            const float x = static_cast<float>(arow(r) - P.my) * (1.f / 16.f);
            float sr = Gr[D], si = Gi[D];
#pragma unroll
            for (int m = D - 1; m >= 0; --m) {
                sr = fmaf(sr, x, Gr[m]);
                si = fmaf(si, x, Gi[m]);
            }
            // origin phase in revolutions, |phs| < 2: within v_sin's own range
            const int p = prow + min(arow(r), P.vy - 1) + tie;
            const float4 du = A.dup[p];
            const float phs = pht - fmaf(koz, du.z, fmaf(koy, du.y, kox * du.x));
            const float sn = __builtin_amdgcn_sinf(phs);
            const float cs = __builtin_amdgcn_cosf(phs);
            acc_re[r] = fmaf(cs, sr, fmaf(-sn, si, acc_re[r]));
            acc_im[r] = fmaf(cs, si, fmaf(sn, sr, acc_im[r]));
        }
        if (++since == POLY_FLUSH || t + 1 == t1) flush();
        if (has_next) {
            float4* dst = reinterpret_cast<float4*>(&tile_buf[buf ^ 1][tid]);
            dst[0] = n0;
            dst[1] = n1;
        }
        __syncthreads();
    }
    if (live && first) {                          // a slice without tiles
#pragma unroll
        for (int r = 0; r < 16; ++r)
            if (arow(r) < P.vy && q < P.vz) out[prow + arow(r)] = make_double2(0.0, 0.0);
    }
}

template <int D, bool SECOND>
hipError_t launch_poly(const MmaArgs& A, int n_slices, int n_frames, hipStream_t st) {
    dim3 grid((A.n_patches + POLY_WAVES - 1) / POLY_WAVES, n_slices, n_frames);
    hipLaunchKernelGGL((k_sum_poly<D, SECOND>), grid, dim3(64 * POLY_WAVES), 0, st, A);
    return hipGetLastError();
}

// Unit of tile t for the moment sums over units and the moment adjoint: the quad, else the pair, whose
// staged radius is within rmax and which covers t, else the tile alone.
// Heads sit at aligned places inside a k-d node and only they carry finite
// radii, so this is the unit the moment sum's walk forms from a node start.
__device__ __forceinline__ void unit_of(int t, int n_tiles, const float* __restrict__ grp_r,
                                        float rmax, int& head, int& size) {
    head = t;
    size = 1;
    if (!(rmax > 0.f)) return;
    for (int h = max(0, t - 3); h <= t; ++h)
        if (h + 3 < n_tiles && grp_r[2 * static_cast<int64_t>(h) + 1] <= rmax) {
            head = h;
            size = 4;
            return;
        }
    for (int h = max(0, t - 1); h <= t; ++h)
        if (h + 1 < n_tiles && grp_r[2 * static_cast<int64_t>(h)] <= rmax) {
            head = h;
            size = 2;
            return;
        }
}

// ------------------------------------------------- moment sum on the matrix cores
// The moment expansion of k_sum_poly with its per-atom work moved to the FP16
// matrix cores.  Per unit of tiles and per patch the moment sum needs
//     C_mn = sum_j a_j T_mn(j),   a_j = |w_j| e^{2 pi i theta_j(patch)},
//     T_mn = P_j^m R_j^n / (m! n!),
// with P_j, R_j the atom offset projected on the patch's row and column
// gradients.  Over patches that share those gradients the T_mn do not depend
// on the patch, and the moments of 16 patches become one real matrix
// product over the unit's atoms: rows (patch, Re or Im of a_j), columns the
// atom's basis functions.  A wave owns a 4 x 4 group of patches and expands
// about the group's mean gradients gy_, gz_.  Each patch's own model differs
// from the group's by
//     eps_j(x, y) = sum_c sigma_jc E_c(x, y),
//     E_c = 2 pi [16 (gy - gy_)_c x + 16 (gz - gz_)_c y + 256 hy_c x^2
//                 + 256 hz_c y^2 + 256 gyz_c x y]
// (its largest value over a patch is a few 1e-3 rad for the largest units;
// the host bounds it), which enters to second order:
//     e^{i (P_ x + R_ y + eps)} = e^{i (P_ x + R_ y)} (1 + i eps - eps^2 / 2) + O(eps^3).
// So each atom's basis row holds (sigma in wavelengths, P_ and R_ in rad)
//     P_^m R_^n                    m + n <= 6   (28 columns)
//     sigma_c P_^m R_^n            m + n <= 4   (45)
//     sigma_c sigma_d {1, P_, R_}   c <= d       (18)
// and a fold per (unit, patch) turns the 91 products into the 28
// coefficients of the patch polynomial S(x, y) = sum M_mn x^m y^n (factorials
// and powers of i included), written to a moment buffer.  A second kernel
// (k_gm_eval) evaluates the polynomials at the pixels with the per-(unit,
// pixel) origin phase, as k_sum_poly does.
//
// Operands: values are split into FP16 halves x = hi + lo (22 significant
// bits) and a product is hi hi + hi lo + lo hi with FP32 accumulation
// (v_mfma_f32_32x32x16_f16: A row = lane % 32, B column = lane % 32, k = 8
// (lane / 32) + 0..7; result row 8 (r / 4) + 4 (lane / 32) + r % 4).  The
// columns sit in three 32-column blocks by the size of their terms: P^m R^n
// (m + n <= 4) and sigma_c P^m R^n in blocks 0 and 1 with the split
// products; P^m R^n (m + n = 5, 6: at most phi^5 / 5! ~ 2e-4 of the sum
// within the degree-6 bound) and sigma_c sigma_d (at most eps^2 / 2 ~ 2e-4)
// in block 2 with hi hi alone, whose rounding (2^-10 of those terms) stays
// below 3e-7 of the sum.  The carrier a_j takes one v_sin per operand value:
// the Re rows read the chunk's atom records with a quarter turn added to the
// phase.  The four waves of a workgroup build the chunk's basis rows into
// shared LDS (a lane per atom, each wave a quarter of the columns), then
// each runs four 16-atom steps of 3 + 3 + 1 products for its patch group.
constexpr int GM_WAVES = 4;                       // patch groups per workgroup
constexpr int GM_GRP = 4;                         // a wave's group: GM_GRP x GM_GRP patches
constexpr int GM_REG = 2 * GM_GRP;                // a workgroup's region: 2 x 2 groups
constexpr int GM_CHUNK = 64;                      // atoms per basis chunk
constexpr int GM_NT = 28;                         // P^m R^n, m + n <= 6
constexpr int GM_N4 = 15;                         // m + n <= 4
constexpr int GM_COL_S = GM_N4;                   // sigma_c P^m R^n: GM_COL_S + 15 c + j
constexpr int GM_COL_T56 = GM_N4 + 3 * GM_N4;     // P^m R^n, m + n = 5 (by m), then 6 (by m)
constexpr int GM_COL_SS = GM_COL_T56 + 13;        // sigma_c sigma_d {1, P, R}: + 3 pair + s
constexpr int GM_NCOL = GM_COL_SS + 18;           // 91
constexpr int GM_NB = 3;                          // 32-column blocks
constexpr int GM_NSPLIT = 64;                     // blocks 0, 1: split products (lo halves kept)
constexpr int GM_BS = 72;                         // halves per basis column: 64 atoms + pad (144 B)
constexpr int GM_CS = 92;                         // floats per product row at a unit's end
static_assert(GM_NCOL <= 32 * GM_NB, "basis columns fit the blocks");
static_assert(GM_CS >= GM_NCOL, "a product row holds every basis column");
static_assert(GM_COL_T56 + 4 == GM_NSPLIT && GM_COL_SS >= GM_NSPLIT,
              "the degree <= 4 and first-order columns fill the split blocks");

__host__ __device__ constexpr int gm_kT(int m, int n) { return 7 * m - m * (m - 1) / 2 + n; }
__host__ __device__ constexpr int gm_k4(int m, int n) { return 5 * m - m * (m - 1) / 2 + n; }
// basis column of P^m R^n, m + n <= 6
__host__ __device__ constexpr int gm_colT(int m, int n) {
    return m + n <= 4 ? gm_k4(m, n) : GM_COL_T56 + (m + n == 5 ? m : 6 + m);
}

struct GmArgs {
    const AtomRec* atoms;
    const TileRec* tiles;
    const FrameRec* frames;
    const double4* u0h;
    const double* theta0;
    const float4* dup;
    const float2* gmma;
    double2* partial;
    float2* mom;              // [head tile - t_begin][frame][patch][GM_REC]
    const uint8_t* heads;     // [tile] unit size of a head tile, else 0
    const float* grp_r;
    float grp_rmax;
    int n_tiles;
    int t_begin;              // this round's unit heads: [t_begin, t_end)
    int t_end;
    int a_tiles_per_slice;    // k_gm_moments' tile slices
    int tiles_per_slice;      // k_gm_eval: the slices of the partials
    int n_slices;
    int n_frames;
    int Ny;
    int Nz;
    int npix;
    int pby;                  // patches along a row / a column
    int pbz;
    int n_patches;
    int ngy;                  // regions of GM_REG x GM_REG patches along a row
    int n_groups;             // regions (one workgroup each)
    int accumulate;           // k_gm_eval adds to the partials (rounds after the first)
    unsigned int* fallbacks;  // (unit, patch) evaluations taken pixel by pixel (counted when not null)
    int pixel_eval;           // evaluate every unit pixel by pixel
    double K;
};

// LDS written by some lanes of a wave and read by others of the same wave:
// the wave's LDS operations complete in order; wait for them (lgkmcnt(0),
// vector memory left in flight) and keep the compiler from moving memory
// operations across.
__device__ __forceinline__ void gm_wave_sync() {
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_wave_barrier();
}

// Unit size of every head tile (0 for the others), from unit_of.
__global__ void k_gm_heads(const float* __restrict__ grp_r, float rmax, int n_tiles,
                           uint8_t* __restrict__ heads) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tiles) return;
    int head, size;
    unit_of(t, n_tiles, grp_r, rmax, head, size);
    heads[t] = head == t ? static_cast<uint8_t>(size) : static_cast<uint8_t>(0);
}

// Walks the unit heads of the tiles [t0, t1) in order, 64 tiles a load: each
// lane takes one tile's unit size, a ballot marks the heads.
struct GmHeadWalk {
    const uint8_t* heads;
    int base, t1, lane, hs;
    uint64_t mask;
    __device__ GmHeadWalk(const uint8_t* h, int t0, int t1_, int lane_)
        : heads(h), base(t0 - 64), t1(t1_), lane(lane_), hs(0), mask(0) {}
    // the next head tile and its unit size; false past the end
    __device__ __forceinline__ bool next(int& t, int& size) {
        while (mask == 0) {
            base += 64;
            if (base >= t1) return false;
            const int tl = base + lane;
            hs = tl < t1 ? heads[tl] : 0;
            mask = __ballot(hs != 0);
        }
        const int i = __builtin_ctzll(mask);
        mask &= mask - 1;
        t = base + i;
        size = __builtin_amdgcn_readlane(hs, i);
        return true;
    }
};

typedef __fp16 halfx2 __attribute__((ext_vector_type(2)));
typedef unsigned int uintx4 __attribute__((ext_vector_type(4)));

// Two operand values into their hi and lo words (value 0 in the low halves):
// hi rounded to nearest, and the residual x - hi by one mixed-precision FMA
// each (the FP16 hi half of the packed word times -1 plus the FP32 value,
// rounded to FP16 into the low or high half of lo).  The round-toward-zero
// packed convert is not used: with it the residual of small values came out
// wrong (values counted twice), which a basis with many small entries showed
// at 3e-4.
__device__ __forceinline__ void gm_split_pair(float x0, float x1, unsigned int& hi,
                                             unsigned int& lo) {
    halfx2 h;
    h[0] = static_cast<__fp16>(x0);
    h[1] = static_cast<__fp16>(x1);
    hi = __builtin_bit_cast(unsigned int, h);
    unsigned int l;
    asm("v_fma_mixlo_f16 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(l) : "v"(hi), "v"(x0));
    asm("v_fma_mixhi_f16 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
        : "+v"(l) : "v"(hi), "v"(x1));
    lo = l;
}

// The hi word alone of a x0 and a x1 (fused, rounded to nearest): for
// operand values small enough that their lo halves are below the error bound.
__device__ __forceinline__ unsigned int gm_prod_hi_pair(float a, float x0, float x1) {
    unsigned int h;
    asm("v_fma_mixlo_f16 %0, %1, %2, 0" : "=v"(h) : "v"(a), "v"(x0));
    asm("v_fma_mixhi_f16 %0, %1, %2, 0" : "+v"(h) : "v"(a), "v"(x1));
    return h;
}

// hi + lo halves of a x0 and a x1 straight from the products (fused, one
// rounding each): hi = f16(a x), lo = f16(a x - hi), four mixed-precision FMAs
// for the pair against two multiplies, a convert and two FMAs.
__device__ __forceinline__ void gm_split_prod_pair(float a, float x0, float x1, unsigned int& hi,
                                                   unsigned int& lo) {
    unsigned int h, l;
    asm("v_fma_mixlo_f16 %0, %1, %2, 0" : "=v"(h) : "v"(a), "v"(x0));
    asm("v_fma_mixhi_f16 %0, %1, %2, 0" : "+v"(h) : "v"(a), "v"(x1));
    asm("v_fma_mixlo_f16 %0, %1, %2, -%3 op_sel_hi:[0,0,1]" : "=v"(l) : "v"(a), "v"(x0), "v"(h));
    asm("v_fma_mixhi_f16 %0, %1, %2, -%3 op_sel:[0,0,1] op_sel_hi:[0,0,1]" : "+v"(l) : "v"(a), "v"(x1), "v"(h));
    hi = h;
    lo = l;
}

// (re, im) times i^k
__device__ __forceinline__ void gm_rot(int k, float& re, float& im) {
    const float r = re, i = im;
    switch (k & 3) {
        case 0: break;
        case 1: re = -i; im = r; break;
        case 2: re = -r; im = -i; break;
        default: re = i; im = -r; break;
    }
}

__device__ __forceinline__ float gm_inv_fact(int m) {
    return m == 0 ? 1.f : m == 1 ? 1.f : m == 2 ? 0.5f : m == 3 ? (1.f / 6.f) :
           m == 4 ? (1.f / 24.f) : m == 5 ? (1.f / 120.f) : (1.f / 720.f);
}

// The fold of patch p = lane % 16 over the four lanes p + 16 j (j = 0..3)
// of a wave: the 91 products (Re in row p, Im in row 16 + p of cb) into the
// 28 polynomial coefficients, with the unit's phase terms at the patch (pht,
// K O_x; K O_y, K O_z) and its species, written to dst when dst is not null.
// Ev[p][5 c .. 5 c + 4] hold the patch's model difference E_c: x, y, x^2,
// y^2, x y coefficients [rad].  Every lane runs the same instructions on its
// own share: lanes j = 0..2 the first-order terms of sigma_j and the
// second-order terms in the form sum_c E_c Y_c, Y_c = -1/2 sum_d E_d W_cd (the
// sum over c <= d of s_cd E_c E_d W_cd with s_cc = -1/2, s_cd = -1 regrouped),
// lane j = 0 the main terms; then a reduce-scatter over the four lanes
// (v_permlane16_swap, v_permlane32_swap) leaves each lane 7 of the 28
// coefficients.  All 64 lanes must take part.
__device__ __forceinline__ void gm_fold4(const float* cb, int lane, const float (*Ev)[15],
                                         float2 ph0, float2 ph1, int species,
                                         float2* __restrict__ dst) {
    constexpr int DM[5] = {1, 0, 2, 0, 1}, DN[5] = {0, 1, 0, 2, 1};   // E_c monomials
    constexpr int YM[9] = {1, 0, 2, 1, 0, 3, 2, 1, 0}, YN[9] = {0, 1, 0, 1, 2, 0, 1, 2, 3};
    auto yidx = [](int m, int n) { return (m + n == 1 ? 0 : (m + n == 2 ? 2 : 5)) + n; };
    const int p = lane & 15, j = lane >> 4;
    const int c = j < 3 ? j : 2;                       // lane j = 3: weight 0
    const float* cr = cb + p * GM_CS;
    const float* ci = cb + (16 + p) * GM_CS;
    float E[3][5], Ec[5];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int u = 0; u < 5; ++u) E[d][u] = Ev[p][5 * d + u];
    const float w1 = j < 3 ? 1.f : 0.f;
#pragma unroll
    for (int u = 0; u < 5; ++u) Ec[u] = Ev[p][5 * c + u] * w1;
    float Mr[GM_MOM], Mi[GM_MOM];
    // main terms (lane j = 0)
    const float wm = j == 0 ? 1.f : 0.f;
#pragma unroll
    for (int m = 0; m <= 6; ++m) {
#pragma unroll
        for (int n = 0; n <= 6 - m; ++n) {
            const int k = gm_kT(m, n), col = gm_colT(m, n);
            const float f = gm_inv_fact(m) * gm_inv_fact(n) * wm;
            float re = cr[col] * f, im = ci[col] * f;
            gm_rot(m + n, re, im);
            Mr[k] = re;
            Mi[k] = im;
        }
    }
    // first order: i E_c(x, y) sum_mn i^{m+n} / (m! n!) C[sigma_c P^m R^n] x^m y^n
    {
        const float* crc = cr + GM_COL_S + GM_N4 * c;
        const float* cic = ci + GM_COL_S + GM_N4 * c;
#pragma unroll
        for (int m = 0; m <= 4; ++m) {
#pragma unroll
            for (int n = 0; n <= 4 - m; ++n) {
                const float f = gm_inv_fact(m) * gm_inv_fact(n);
                float vr = crc[gm_k4(m, n)] * f, vi = cic[gm_k4(m, n)] * f;
                gm_rot(m + n + 1, vr, vi);
#pragma unroll
                for (int u = 0; u < 5; ++u) {
                    const int k = gm_kT(m + DM[u], n + DN[u]);
                    Mr[k] = fmaf(Ec[u], vr, Mr[k]);
                    Mi[k] = fmaf(Ec[u], vi, Mi[k]);
                }
            }
        }
    }
    // second order: E_c Y_c, Y_c = -1/2 sum_d E_d (C[s_c s_d] + i C[s_c s_d P] x + i C[s_c s_d R] y)
    {
        float Yr[9], Yi[9];
#pragma unroll
        for (int v = 0; v < 9; ++v) { Yr[v] = 0.f; Yi[v] = 0.f; }
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int a = c < d ? c : d, b = c < d ? d : c;
            const int pair = a == 0 ? b : (a == 1 ? 2 + b : 5);
            const int col = GM_COL_SS + 3 * pair;
            const float wr[3] = {cr[col], -ci[col + 1], -ci[col + 2]};
            const float wi[3] = {ci[col], cr[col + 1], cr[col + 2]};
#pragma unroll
            for (int u = 0; u < 5; ++u)
#pragma unroll
                for (int t = 0; t < 3; ++t) {
                    const int y = yidx(DM[u] + (t == 1 ? 1 : 0), DN[u] + (t == 2 ? 1 : 0));
                    Yr[y] = fmaf(E[d][u], wr[t], Yr[y]);
                    Yi[y] = fmaf(E[d][u], wi[t], Yi[y]);
                }
        }
#pragma unroll
        for (int u = 0; u < 5; ++u) {
            const float e = -0.5f * Ec[u];
#pragma unroll
            for (int v = 0; v < 9; ++v) {
                const int k = gm_kT(DM[u] + YM[v], DN[u] + YN[v]);
                Mr[k] = fmaf(e, Yr[v], Mr[k]);
                Mi[k] = fmaf(e, Yi[v], Mi[k]);
            }
        }
    }
    // reduce-scatter over the patch's four lanes: values v[2 k] = Mr[k],
    // v[2 k + 1] = Mi[k]; rows 0, 1 (2, 3) of 16 lanes pair up first, then
    // the two halves; lane j ends with v[14 q .. 14 q + 13], q = (j >> 1) | 2 (j & 1)
    float h[28];
#pragma unroll
    for (int i = 0; i < 28; ++i) {
        const float x = (i & 1) ? Mi[i >> 1] : Mr[i >> 1];
        const float y = ((28 + i) & 1) ? Mi[(28 + i) >> 1] : Mr[(28 + i) >> 1];
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(y), false, false);
        h[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    float o[14];
#pragma unroll
    for (int i = 0; i < 14; ++i) {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(h[i]), __float_as_uint(h[14 + i]),
                                                        false, false);
        o[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    if (dst) {
        const int qb = 7 * ((j >> 1) | ((j & 1) << 1));
#pragma unroll
        for (int i = 0; i < 7; ++i) dst[qb + i] = make_float2(o[2 * i], o[2 * i + 1]);
        if (j == 0) {
            dst[GM_MOM] = ph0;
            dst[GM_MOM + 1] = ph1;
            dst[GM_MOM + 2] = make_float2(__int_as_float(species), 0.f);
            dst[GM_MOM + 3] = make_float2(0.f, 0.f);
        }
    }
}

// Phase terms of a unit with origin (uox, uoy, uoz) [m, sample frame] at a
// patch's middle pixel (the direction um, Theta0 th_mid, direction-model
// centre vc0): the carrier direction v (sample frame, first order in the
// origin's offset), K times the lab origin and the origin phase at the
// middle pixel, as k_sum_poly forms them.
struct GmUnitPhase {
    float vx, vy, vz, kox, koy, koz, pht;
};
__device__ __forceinline__ GmUnitPhase gm_unit_phase(const FrameRec& F, double uox, double uoy,
                                                     double uoz, const double4& um,
                                                     double th_mid, const double (&vc0)[3],
                                                     double K) {
    const double Ox = fma(F.R[0], uox, fma(F.R[1], uoy, fma(F.R[2], uoz, F.T[0])));
    const double Oy = fma(F.R[3], uox, fma(F.R[4], uoy, fma(F.R[5], uoz, F.T[1])));
    const double Oz = fma(F.R[6], uox, fma(F.R[7], uoy, fma(F.R[8], uoz, F.T[2])));
    const double O2 = Ox * Ox + Oy * Oy + Oz * Oz;
    const double uo = um.x * Ox + um.y * Oy + um.z * Oz;
    double ph_t = th_mid + K * (Ox - uo + (O2 - uo * uo) * um.w);
    ph_t -= floor(ph_t);
    const double h2 = 2.0 * um.w;
    const double cxl = (Ox - uo * um.x) * h2, cyl = (Oy - uo * um.y) * h2,
                 czl = (Oz - uo * um.z) * h2;
    GmUnitPhase r;
    r.vx = static_cast<float>(vc0[0] + (F.R[0] * cxl + F.R[3] * cyl + F.R[6] * czl));
    r.vy = static_cast<float>(vc0[1] + (F.R[1] * cxl + F.R[4] * cyl + F.R[7] * czl));
    r.vz = static_cast<float>(vc0[2] + (F.R[2] * cxl + F.R[5] * cyl + F.R[8] * czl));
    r.kox = static_cast<float>(K * Ox);
    r.koy = static_cast<float>(K * Oy);
    r.koz = static_cast<float>(K * Oz);
    r.pht = static_cast<float>(ph_t);
    return r;
}

// Origin of the unit headed by tile t with gsz tiles.
__device__ __forceinline__ void gm_unit_origin(const TileRec* __restrict__ tiles, int t, int gsz,
                                               const TileRec& tr, double& ox, double& oy,
                                               double& oz) {
    ox = tr.ox;
    oy = tr.oy;
    oz = tr.oz;
    if (gsz > 1) {
        const TileRec b1 = tiles[t + 1];
        const TileRec b2 = tiles[gsz == 4 ? t + 2 : t];
        const TileRec b3 = tiles[gsz == 4 ? t + 3 : t];
        ox = unit_origin(gsz, tr.ox, b1.ox, b2.ox, b3.ox);
        oy = unit_origin(gsz, tr.oy, b1.oy, b2.oy, b3.oy);
        oz = unit_origin(gsz, tr.oz, b1.oz, b2.oz, b3.oz);
    }
}

// The patch's direction model (sample frame) as k_sum_poly fits it: centre
// vc0 and per pixel the row and column slopes g, curvatures h and the mixed
// term gyz.
__device__ __forceinline__ void gm_patch_model(const FrameRec& F, const double4* __restrict__ u0h,
                                               const PatchGeom& P, int Ny, double (&vc0)[3],
                                               double (&gy)[3], double (&hy)[3], double (&gz)[3],
                                               double (&hz)[3], double (&gyz)[3]) {
    const int pmid = (P.z0 + P.mz) * Ny + P.y0 + P.my;
    const int pc = pmid - P.my - P.mz * Ny;
    const int samp[9] = {pmid, pmid - P.my, pmid - P.my + P.vy - 1, pmid - P.mz * Ny,
                         pmid + (P.vz - 1 - P.mz) * Ny, pc, pc + P.vy - 1,
                         pc + (P.vz - 1) * Ny, pc + P.vy - 1 + (P.vz - 1) * Ny};
    double vs[9][3];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        const double4 u = u0h[samp[k]];
        const double wx = 1.0 - u.x, wy = -u.y, wz = -u.z;
        vs[k][0] = F.R[0] * wx + F.R[3] * wy + F.R[6] * wz;
        vs[k][1] = F.R[1] * wx + F.R[4] * wy + F.R[7] * wz;
        vs[k][2] = F.R[2] * wx + F.R[5] * wy + F.R[8] * wz;
    }
    const double area = static_cast<double>(P.vy - 1) * static_cast<double>(P.vz - 1);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        vc0[d] = vs[0][d];
        quad_fit(vs[1][d] - vs[0][d], vs[2][d] - vs[0][d], -P.my, P.vy - 1 - P.my, gy[d], hy[d]);
        quad_fit(vs[3][d] - vs[0][d], vs[4][d] - vs[0][d], -P.mz, P.vz - 1 - P.mz, gz[d], hz[d]);
        gyz[d] = area > 0.0 ? (vs[8][d] - vs[7][d] - vs[6][d] + vs[5][d]) / area : 0.0;
    }
}

template <bool SECOND>
__global__ __launch_bounds__(64 * GM_WAVES) __attribute__((amdgpu_waves_per_eu(2)))
void k_gm_moments(const GmArgs A) {
    // Two chunk buffers, each: the basis hi halves [column][atom position],
    // then the lo halves of blocks 0 and 1; and a third region.  At a unit's
    // end waves 0, 1 put their products for the fold into the buffer the
    // last chunk read, waves 2, 3 into the third region, while the other
    // buffer holds the next unit's first chunk.
    constexpr int SB_H = GM_NB * 32 * GM_BS * 2, SB_L = GM_NSPLIT * GM_BS * 2;
    constexpr int SB_P = 2 * 32 * GM_CS * 4;
    constexpr int SB = SB_H + SB_L > SB_P ? SB_H + SB_L : SB_P;
    static_assert(SB % 16 == 0 && GM_WAVES == 4, "the regions stay 16-byte aligned, two waves to each");
    __shared__ __align__(16) unsigned char smem[3 * SB];
    __shared__ float sev[GM_WAVES][16][15];                      // each patch's E_c coefficients
    __shared__ double sgs[GM_WAVES][8];                          // the waves' gradient sums
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int frame = blockIdx.z, slice = blockIdx.y;
    const FrameRec& F = A.frames[frame];
    const int g = lane >> 5, q = lane & 31, pl = lane & 15;
    // this workgroup's region of GM_REG x GM_REG patches, this wave's 4 x 4 group in it
    const int rgn = blockIdx.x;
    const int gy0 = (rgn % A.ngy) * GM_REG + GM_GRP * (wave & 1);
    const int gz0 = (rgn / A.ngy) * GM_REG + GM_GRP * (wave >> 1);
    const int ppy = gy0 + (pl & 3), ppz = gz0 + (pl >> 2);
    const bool plive = ppy < A.pby && ppz < A.pbz;
    const PatchGeom P = patch_of(min(ppy, A.pby - 1), min(ppz, A.pbz - 1), A.Ny, A.Nz);
    const int pmid = (P.z0 + P.mz) * A.Ny + P.y0 + P.my;

    // this lane's patch model; the region's mean gradients over its live
    // patches (lanes 0..15 of each wave hold its patches once)
    double vc0[3], gyv[3], hyv[3], gzv[3], hzv[3], gyzv[3];
    gm_patch_model(F, A.u0h, P, A.Ny, vc0, gyv, hyv, gzv, hzv, gyzv);
    {
        double sg[7];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            sg[d] = plive ? gyv[d] : 0.0;
            sg[3 + d] = plive ? gzv[d] : 0.0;
        }
        sg[6] = plive ? 1.0 : 0.0;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1)
#pragma unroll
            for (int k = 0; k < 7; ++k) sg[k] += __shfl_xor(sg[k], o, 64);
        if (lane == 0)
#pragma unroll
            for (int k = 0; k < 7; ++k) sgs[wave][k] = sg[k];
        // (the basis pad columns 91..95 are not written: their products land
        // in result columns that are never read)
    }
    __syncthreads();
    double mg[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int w = 0; w < GM_WAVES; ++w)
#pragma unroll
        for (int k = 0; k < 7; ++k) mg[k] += sgs[w][k];
    const double inv_n = 1.0 / mg[6];             // the region's first patch is live
    constexpr double TP16 = 2.0 * PI * 16.0, TP256 = 2.0 * PI * 256.0;
    float gby[3], gbz[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double my = mg[d] * inv_n, mz = mg[3 + d] * inv_n;
        gby[d] = static_cast<float>(TP16 * my);
        gbz[d] = static_cast<float>(TP16 * mz);
        if (lane < 16) {
            float* ev = sev[wave][lane] + 5 * d;
            ev[0] = static_cast<float>(TP16 * (gyv[d] - my));
            ev[1] = static_cast<float>(TP16 * (gzv[d] - mz));
            ev[2] = static_cast<float>(TP256 * hyv[d]);
            ev[3] = static_cast<float>(TP256 * hzv[d]);
            ev[4] = static_cast<float>(TP256 * gyzv[d]);
        }
    }
    const double4 um = A.u0h[pmid];
    const double th_mid = A.theta0[pmid];
    const float xc0 = static_cast<float>(F.xs[0] - vc0[0]);
    const float xc1 = static_cast<float>(F.xs[1] - vc0[1]);
    const float xc2 = static_cast<float>(F.xs[2] - vc0[2]);
    const float hh = SECOND ? static_cast<float>(um.w / A.K) : 0.f;

    const int t0 = A.t_begin + slice * A.a_tiles_per_slice;
    const int t1 = min(A.t_end, t0 + A.a_tiles_per_slice);
    GmHeadWalk walk(A.heads, t0, t1, lane);      // every wave walks the same units
    // a unit: head tile t, gsz tiles, its origin (a unit of 2 repeats its
    // second tile in the mean) and its head's species
    struct UnitAt {
        int t, gsz, species;
        double ox, oy, oz;
    };
    auto unit_at = [&](int t, int gsz) {
        TileRec trk[4];
        trk[0] = A.tiles[t];
#pragma unroll
        for (int k = 1; k < 4; ++k) trk[k] = A.tiles[t + min(k, gsz - 1)];
        UnitAt u;
        u.t = t;
        u.gsz = gsz;
        u.species = trk[0].species;
        u.ox = trk[0].ox;
        u.oy = trk[0].oy;
        u.oz = trk[0].oz;
        if (gsz > 1) {
            u.ox = unit_origin(gsz, trk[0].ox, trk[1].ox, trk[2].ox, trk[3].ox);
            u.oy = unit_origin(gsz, trk[0].oy, trk[1].oy, trk[2].oy, trk[3].oy);
            u.oz = unit_origin(gsz, trk[0].oz, trk[1].oz, trk[2].oz, trk[3].oz);
        }
        return u;
    };
    // a chunk's place (tile tl, ti in its unit, offset co) with its tile's
    // count and origin offset from the unit origin [wavelengths]
    struct ChunkAt {
        int tl, ti, co, cnt;
        float dx, dy, dz;
    };
    auto at_tile = [&](const UnitAt& u, int ti) {
        const TileRec r = A.tiles[u.t + ti];
        ChunkAt c;
        c.tl = u.t + ti;
        c.ti = ti;
        c.co = 0;
        c.cnt = r.count & 0xFFFF;
        c.dx = static_cast<float>((r.ox - u.ox) * A.K);
        c.dy = static_cast<float>((r.oy - u.oy) * A.K);
        c.dz = static_cast<float>((r.oz - u.oz) * A.K);
        return c;
    };
    // the chunk after c in unit u, false past the unit's last
    auto advance = [&](const UnitAt& u, ChunkAt& c) {
        c.co += GM_CHUNK;
        if (c.co >= c.cnt) {
            if (c.ti + 1 >= u.gsz) return false;
            c = at_tile(u, c.ti + 1);
        }
        return true;
    };
    auto rec = [&](const ChunkAt& c) {
        return A.atoms[static_cast<int64_t>(c.tl) * TILE + c.co + lane];
    };
    GmUnitPhase U;
    floatx16 c0, c1, c2;

    // A chunk's basis rows, each value times the atom's |w| (the carrier
    // operand is then a plain sine): wave w writes the columns f_w P^m
    // R^n (m + n <= 4), f = 1, sigma_x, sigma_y, sigma_z, in four parts
    // placed between the previous chunk's product steps, and then eight of
    // the rest (records past a tile's count are zero: weight 0).
    struct Rows {
        float sx, sy, sz, f, Pb, Rb, ph;
        float pw[7], rw[7];
    };
    auto rows_of = [&](const AtomRec& at, const ChunkAt& ca) {
        Rows r;
        r.sx = at.sx + ca.dx;
        r.sy = at.sy + ca.dy;
        r.sz = at.sz + ca.dz;
        r.ph = at.ph;
        r.Pb = fmaf(gby[2], r.sz, fmaf(gby[1], r.sy, gby[0] * r.sx));
        r.Rb = fmaf(gbz[2], r.sz, fmaf(gbz[1], r.sy, gbz[0] * r.sx));
        r.pw[0] = at.mag;
        r.rw[0] = 1.f;
#pragma unroll
        for (int m = 1; m <= 6; ++m) {
            r.pw[m] = r.pw[m - 1] * r.Pb;
            r.rw[m] = r.rw[m - 1] * r.Rb;
        }
        r.f = wave == 0 ? 1.f : (wave == 1 ? r.sx : (wave == 2 ? r.sy : r.sz));
        return r;
    };
    // Atom a of a chunk sits at position a with bits 2 and 3 swapped: the
    // order in which the phase products leave the atoms in each lane.
    const int apos = (lane & ~12) | ((lane & 4) << 1) | ((lane & 8) >> 1);
    auto put_split = [&](_Float16* bh, _Float16* bl, int col, float v) {   // col < GM_NSPLIT
        const _Float16 h = static_cast<_Float16>(v);
        bh[col * GM_BS + apos] = h;
        bl[col * GM_BS + apos] = static_cast<_Float16>(v - static_cast<float>(h));
    };
    // Two columns' values of this lane's atom into LDS as 32-bit words of
    // two atoms: lanes 2k and 2k + 1 swap one value (DPP quad_perm
    // [1, 0, 3, 2]), then lane 2k stores column A's word of atoms 2k, 2k + 1
    // and lane 2k + 1 column B's (hi and lo words for split columns, else
    // the hi word rounded to nearest).
    auto put_pair = [&](_Float16* bh, _Float16* bl, int colA, float vA, int colB, float vB,
                        bool split) {
        const bool odd = (lane & 1) != 0;
        const float x = odd ? vA : vB;
        const float y = __int_as_float(
            __builtin_amdgcn_update_dpp(0, __float_as_int(x), 0xB1, 0xF, 0xF, false));
        const float first = odd ? y : vA, second = odd ? vB : y;
        const int w = (odd ? colB : colA) * (GM_BS / 2) + (apos >> 1);
        if (split) {
            unsigned int hi, lo;
            gm_split_pair(first, second, hi, lo);
            reinterpret_cast<unsigned int*>(bh)[w] = hi;
            reinterpret_cast<unsigned int*>(bl)[w] = lo;
        } else {
            halfx2 h;
            h[0] = static_cast<__fp16>(first);
            h[1] = static_cast<__fp16>(second);
            reinterpret_cast<unsigned int*>(bh)[w] = __builtin_bit_cast(unsigned int, h);
        }
    };
    // part 0..3 of the main columns (the 15 of m + n <= 4 in order, four
    // to a part, the last alone) into buffer B
    auto build_main = [&](int b, const Rows& r, int part) {
        constexpr int KM[15] = {0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 3, 3, 4};
        constexpr int KN[15] = {0, 1, 2, 3, 4, 0, 1, 2, 3, 0, 1, 2, 0, 1, 0};
        _Float16* const bh = reinterpret_cast<_Float16*>(smem + b * SB);
        _Float16* const bl = reinterpret_cast<_Float16*>(smem + b * SB + SB_H);
        const int col0 = wave == 0 ? 0 : GM_COL_S + GM_N4 * (wave - 1);
        auto val = [&](int i) { return r.f * (r.pw[KM[i]] * r.rw[KN[i]]); };
#pragma unroll
        for (int i = 4 * part; i < 4 * part + 4 && i < 15; i += 2) {
            if (i + 1 < 15) put_pair(bh, bl, col0 + i, val(i), col0 + i + 1, val(i + 1), true);
            else put_split(bh, bl, col0 + i, val(i));
        }
    };
    // P^m R^n, m + n = 5, 6 (13) and sigma_c sigma_d {1, P, R} (18):
    // waves 0..3 take eight each (the last one a pad column), in pairs
    auto build_rest = [&](int b, const Rows& r) {
        _Float16* const bh = reinterpret_cast<_Float16*>(smem + b * SB);
        _Float16* const bl = reinterpret_cast<_Float16*>(smem + b * SB + SB_H);
        auto pair = [&](int ca, float va, int cb, float vb) {   // both split or both not
            put_pair(bh, bl, ca, va, cb, vb, ca < GM_NSPLIT);
        };
        const float ms = r.pw[0];
        const float ss[6] = {ms * r.sx * r.sx, ms * r.sx * r.sy, ms * r.sx * r.sz,
                             ms * r.sy * r.sy, ms * r.sy * r.sz, ms * r.sz * r.sz};
        const float s3[3] = {1.f, r.Pb, r.Rb};
        const float* pw = r.pw;
        const float* rw = r.rw;
        auto sig = [&](int i) { return ss[i / 3] * s3[i % 3]; };
        static_assert(gm_colT(0, 5) + 4 == GM_NSPLIT, "split and hi-only pairs of wave 0");
        if (wave == 0) {
            pair(gm_colT(0, 5), pw[0] * rw[5], gm_colT(1, 4), pw[1] * rw[4]);
            pair(gm_colT(2, 3), pw[2] * rw[3], gm_colT(3, 2), pw[3] * rw[2]);
            pair(gm_colT(4, 1), pw[4] * rw[1], gm_colT(5, 0), pw[5] * rw[0]);
            pair(gm_colT(0, 6), pw[0] * rw[6], gm_colT(1, 5), pw[1] * rw[5]);
        } else if (wave == 1) {
            pair(gm_colT(2, 4), pw[2] * rw[4], gm_colT(3, 3), pw[3] * rw[3]);
            pair(gm_colT(4, 2), pw[4] * rw[2], gm_colT(5, 1), pw[5] * rw[1]);
            pair(gm_colT(6, 0), pw[6] * rw[0], GM_COL_SS + 0, sig(0));
            pair(GM_COL_SS + 1, sig(1), GM_COL_SS + 2, sig(2));
        } else if (wave == 2) {
#pragma unroll
            for (int i = 3; i < 11; i += 2) pair(GM_COL_SS + i, sig(i), GM_COL_SS + i + 1, sig(i + 1));
        } else {
#pragma unroll
            for (int i = 11; i < 17; i += 2) pair(GM_COL_SS + i, sig(i), GM_COL_SS + i + 1, sig(i + 1));
            pair(GM_COL_SS + 17, sig(17), GM_NCOL, 0.f);   // with the first pad column
        }
    };
    // The chunk's phases th(atom, row) = v . s + ph (+ 1/4 on the Re rows
    // q < 16; with SECOND + hh (|s|^2 - (xc . s)^2)) as products on the
    // matrix cores, 32 atoms x 32 rows from K = 2 (FP32, one per atom
    // pair of operands): lane l holds A[atom l & 31][k = l >> 5] and
    // B[k = l >> 5][row l & 31], so the operands are the atom records'
    // (x, y) and (z, ph) pairs with lanes 32..63 holding the second of
    // each, for atoms 0..31 and 32..63 (two v_permlane32_swap), against
    // the lane patch's (vx, vy), (vz, 1).  The result leaves lane l the
    // phases of row l & 31 at atoms 8 (r >> 2) + 4 (l >> 5) + (r & 3),
    // r = 0..15: steps 0, 1 (2, 3) take r = 0..7, 8..15 of the first
    // (second) 32 atoms, the atoms at their basis positions.
    struct PhOps {
        float xy0, zp0, s0, xy1, zp1, s1;
    };
    auto ph_ops = [&](const Rows& r) {
        PhOps o;
        const auto xy = __builtin_amdgcn_permlane32_swap(__float_as_uint(r.sx), __float_as_uint(r.sy),
                                                         false, false);
        const auto zp = __builtin_amdgcn_permlane32_swap(__float_as_uint(r.sz), __float_as_uint(r.ph),
                                                         false, false);
        o.xy0 = __uint_as_float(xy[0]);
        o.xy1 = __uint_as_float(xy[1]);
        o.zp0 = __uint_as_float(zp[0]);
        o.zp1 = __uint_as_float(zp[1]);
        o.s0 = o.s1 = 0.f;
        if (SECOND) {
            const float s2 = fmaf(r.sz, r.sz, fmaf(r.sy, r.sy, r.sx * r.sx));
            const auto ss = __builtin_amdgcn_permlane32_swap(__float_as_uint(s2), 0u, false, false);
            o.s0 = __uint_as_float(ss[0]);
            o.s1 = __uint_as_float(ss[1]);
        }
        return o;
    };
    float pb0 = 0.f, pb1 = 0.f;                      // the unit's: g == 0 ? (vx, vz) : (vy, 1)
    const float pq = q < 16 ? 0.25f : 0.f;
    const float pw0 = g == 0 ? xc0 : xc1, pw1 = g == 0 ? xc2 : 0.f, ps = g == 0 ? hh : 0.f;
    auto phases = [&](float xy, float zp, float s2) {
        floatx16 c;
#pragma unroll
        for (int r = 0; r < 16; ++r) c[r] = pq;
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(xy, pb0, c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_32x32x2f32(zp, pb1, c, 0, 0, 0);
        if (SECOND) {
            c = __builtin_amdgcn_mfma_f32_32x32x2f32(s2, ps, c, 0, 0, 0);
            floatx16 w;
#pragma unroll
            for (int r = 0; r < 16; ++r) w[r] = 0.f;
            w = __builtin_amdgcn_mfma_f32_32x32x2f32(xy, pw0, w, 0, 0, 0);
            w = __builtin_amdgcn_mfma_f32_32x32x2f32(zp, pw1, w, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) c[r] = fmaf(-hh * w[r], w[r], c[r]);
        }
        return c;
    };
    // Product step s (16 atoms) of the chunk in buffer B: the carrier (one
    // sine per operand value, from the phases th) against the three blocks.
    auto product_step = [&](int b, int s, const floatx16& th) {
        const _Float16* const bh = reinterpret_cast<const _Float16*>(smem + b * SB);
        const _Float16* const bl = reinterpret_cast<const _Float16*>(smem + b * SB + SB_H);
        const int r0 = 8 * (s & 1);
        uintx4 hw, lw;
#pragma unroll
        for (int u2 = 0; u2 < 4; ++u2) {
            unsigned int hv, lv;
            gm_split_pair(__builtin_amdgcn_sinf(th[r0 + 2 * u2]), __builtin_amdgcn_sinf(th[r0 + 2 * u2 + 1]),
                          hv, lv);
            hw[u2] = hv;
            lw[u2] = lv;
        }
        const halfx8 ah = __builtin_bit_cast(halfx8, hw);
        const halfx8 al = __builtin_bit_cast(halfx8, lw);
        const int boff = 16 * s + 8 * g;
        const halfx8 bh0 = *reinterpret_cast<const halfx8*>(bh + q * GM_BS + boff);
        const halfx8 bl0 = *reinterpret_cast<const halfx8*>(bl + q * GM_BS + boff);
        const halfx8 bh1 = *reinterpret_cast<const halfx8*>(bh + (32 + q) * GM_BS + boff);
        const halfx8 bl1 = *reinterpret_cast<const halfx8*>(bl + (32 + q) * GM_BS + boff);
        const halfx8 bh2 = *reinterpret_cast<const halfx8*>(bh + (64 + q) * GM_BS + boff);
        c0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh0, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh1, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh2, c2, 0, 0, 0);
        c0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl0, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl1, c1, 0, 0, 0);
        c0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh0, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh1, c1, 0, 0, 0);
    };

    // The units in order.  A unit's 64-atom chunks go in order through the
    // two chunk buffers: the next chunk's basis goes into the other buffer
    // between the current chunk's product steps, one barrier per chunk; each
    // chunk's records are loaded a chunk ahead of its basis (every wave loads
    // all 64: a lane builds one atom's share).  During a unit's last chunk
    // the next unit's first chunk is built, so only the first unit builds
    // one on its own.  The products at a unit's end go to the buffer its last
    // chunk read (waves 0, 1) and the third region (waves 2, 3) for the fold.
    int t, gsz;
    if (!walk.next(t, gsz)) return;                 // (uniform over the workgroup)
    UnitAt cur = unit_at(t, gsz);
    ChunkAt first = at_tile(cur, 0);
    ChunkAt cn = first;
    bool more = advance(cur, cn);                   // cn: the chunk after the first
    AtomRec nxt;                                    // the record of more ? cn : fb
    AtomRec pre = rec(cn);                          // cn's record, loaded ahead (when more)
    auto unit_phase = [&]() {
        U = gm_unit_phase(F, cur.ox, cur.oy, cur.oz, um, th_mid, vc0, A.K);
        pb0 = g == 0 ? U.vx : U.vy;
        pb1 = g == 0 ? U.vz : 1.f;
    };
    unit_phase();
    PhOps pa;
    floatx16 thA;                                  // the current chunk's first 32 atoms' phases
    {
        const Rows r0 = rows_of(rec(first), first);
#pragma unroll
        for (int part = 0; part < 4; ++part) build_main(0, r0, part);
        build_rest(0, r0);
        pa = ph_ops(r0);
        thA = phases(pa.xy0, pa.zp0, pa.s0);
    }
    __syncthreads();
    int b = 0;                                     // the current chunk's buffer
    for (;;) {
        // the next unit (its first chunk is built during this unit's last),
        // else this unit's first chunk again into the buffer nothing reads
        int tn, gn;
        const bool have_next = walk.next(tn, gn);
        UnitAt nu = cur;
        ChunkAt fb = first;
        if (have_next) {
            nu = unit_at(tn, gn);
            fb = at_tile(nu, 0);
        }
        ChunkAt cn1 = fb;                              // the next unit's second chunk
        const bool more1 = have_next && advance(nu, cn1);
        nxt = more ? pre : rec(fb);
#pragma unroll
        for (int r = 0; r < 16; ++r) { c0[r] = 0.f; c1[r] = 0.f; c2[r] = 0.f; }
        // one chunk: products from buffer b, the next chunk's rows into the other
        auto chunk = [&]() {
            ChunkAt c2n = cn;
            const bool more2 = more && advance(cur, c2n);
            const AtomRec nxt2 = rec(more2 ? c2n : fb);
            const Rows r = rows_of(nxt, more ? cn : fb);
            if (!more && more1) pre = rec(cn1);    // the next unit's, a unit ahead
            floatx16 thB;                          // the second 32 atoms' phases
#pragma unroll
            for (int s = 0; s < GM_CHUNK / 16; ++s) {
                if (s == 1) thB = phases(pa.xy1, pa.zp1, pa.s1);
                product_step(b, s, s < 2 ? thA : thB);
                build_main(b ^ 1, r, s);
            }
            build_rest(b ^ 1, r);
            pa = ph_ops(r);
            if (more) thA = phases(pa.xy0, pa.zp0, pa.s0);   // (the next unit's: after its phase terms)
            __syncthreads();                       // buffer b is free, the other complete
            const bool had = more;
            cn = c2n;
            more = more2;
            nxt = nxt2;
            return had;
        };
        while (chunk()) b ^= 1;
        // the products into this wave's LDS (rows: Re of patches 0..15, then
        // Im), then the fold
        float* const cbuf = reinterpret_cast<float*>(smem + (wave < 2 ? b : 2) * SB) + (wave & 1) * 32 * GM_CS;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 8 * (r >> 2) + 4 * g + (r & 3);
            cbuf[row * GM_CS + q] = c0[r];
            cbuf[row * GM_CS + 32 + q] = c1[r];
            if (64 + q < GM_NCOL) cbuf[row * GM_CS + 64 + q] = c2[r];   // not the pad columns
        }
        gm_wave_sync();
        {
            const int patch = ppz * A.pby + ppy;
            float2* dst = plive ? A.mom + ((static_cast<int64_t>(cur.t - A.t_begin) * A.n_frames + frame) *
                                               A.n_patches + patch) * GM_REC
                                : nullptr;
            gm_fold4(cbuf, lane, sev[wave], make_float2(U.pht, U.kox), make_float2(U.koy, U.koz),
                     cur.species, dst);
        }
        __syncthreads();                           // buffer b takes the next unit's second chunk
        if (!have_next) break;
        cur = nu;
        first = fb;
        cn = cn1;
        more = more1;
        unit_phase();
        thA = phases(pa.xy0, pa.zp0, pa.s0);
        b ^= 1;
    }
}

template <bool SECOND>
hipError_t launch_gm_moments(const GmArgs& A, int slices, hipStream_t st) {
    dim3 grid(A.n_groups, slices, A.n_frames);
    hipLaunchKernelGGL((k_gm_moments<SECOND>), grid, dim3(64 * GM_WAVES), 0, st, A);
    return hipGetLastError();
}

// Evaluation of the moment buffer: one wave per patch walks the unit heads
// of its slice (within this round's [t_begin, t_end)) and accumulates each
// unit's polynomial P_u(x, y) = sum M_mn x^m y^n times its origin phase
//     e^{2 pi i (pht_u - K O_u . du(a, b))}
// at the patch's 32 x 32 pixels (du = u_p - u_mid, a along the row, b down
// the column).  With a reference origin O_s per slice (its first unit's) the
// phase splits into e^{-2 pi i K O_s . du}, applied exactly per pixel when
// the sums are flushed, and the unit's residual K dO . du, dO = O_u - O_s,
// whose du is modelled as dr(a) + dc(b) + (a - my)(b - mz) h: the exact
// offsets along the patch's middle row and column and a bilinear cross term
// (h from the patch corners).  Then, to first order in the cross term,
//     S_u(a, b) = alpha(a) beta(b) [P_u - 256 i kappa x y P_u],
//     alpha = e^{-2 pi i K dO . dr(a)},  beta = e^{2 pi i pht_u} e^{-2 pi i K dO . dc(b)},
//     kappa = 2 pi K dO . h,
// a sum of 14 products L(a, k) N(k, b) with L = alpha x^m (main) or alpha
// x^{m+1} (cross), N = beta G_m(y) or -256 i kappa y beta G_m(y), G_m(y) =
// sum_n M_mn y^n: one complex 32 x 32 x 16 product per unit, which the
// matrix cores accumulate over the units in the pixels' running sums (FP16
// split operands, 12 products).  Each half of the wave holds the k slots of
// four powers m (so a lane forms four G_m and four T = beta G_m, not seven).  The model's error, K |dO| times the largest
// deviation of the lane pixels' du from it (measured per patch), must stay
// within 1e-6 revolutions and the cross term's second order within 1e-6;
// a unit that fails either is evaluated pixel by pixel with its exact
// phase (Horner's rule and one sincos per pixel), into the same sums.  The
// running sums go, times the pixel table of their species and the slice's
// reference phase, into the slice's FP64 partial at a species change, every
// POLY_FLUSH units and at the end.
__global__ __launch_bounds__(64 * POLY_WAVES) __attribute__((amdgpu_waves_per_eu(3)))
void k_gm_eval(const GmArgs A) {
    constexpr int D = 6;
    constexpr float TWO_PI_F = 6.283185307179586f;
    // a record's copy in LDS: power m's moments M_mn from GM_PB[m], each run
    // padded with zeros to the Horner length of the power's k slot (so the
    // shorter half of a slot reads zeros past its own terms), complex entry
    // 31 zero for the pad slot; the record's extra float2 from float GM_EVX
    constexpr int GM_PB[8] = {0, 7, 14, 19, 24, 27, 30, 31};
    constexpr int GM_EVX = 64;
    __shared__ __align__(16) float srec[POLY_WAVES][GM_EVX + 2 * (GM_REC - GM_MOM)];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int patch = blockIdx.x * POLY_WAVES + wave;
    const bool live = patch < A.n_patches;
    const int frame = blockIdx.z;
    const int slice = blockIdx.y;
    const int s0 = slice * A.tiles_per_slice;
    const int t0 = max(A.t_begin, s0);
    const int t1 = min(min(A.t_end, A.n_tiles), s0 + A.tiles_per_slice);
    const int pc = live ? patch : 0;
    const PatchGeom P = patch_of(pc % A.pby, pc / A.pby, A.Ny, A.Nz);
    const int Ny = A.Ny;
    const int g = lane >> 5;
    const int q = lane & 31;
    const int prow = (P.z0 + min(q, P.vz - 1)) * Ny + P.y0;
    auto arow = [&](int r) { return 8 * (r >> 2) + 4 * g + (r & 3); };
    auto pix = [&](int r) { return prow + min(arow(r), P.vy - 1); };
    float* const rw = srec[wave];
    int dpos;                                          // this lane's float of the record, in rw
    {
        const int k = lane >> 1;
        int m = 0;
        while (m < 6 && gm_kT(m + 1, 0) <= k) ++m;
        dpos = k < GM_MOM ? 2 * (GM_PB[m] + k - gm_kT(m, 0)) + (lane & 1) : GM_EVX + lane - 2 * GM_MOM;
        const int zc = lane < 2 ? 13 : (lane < 4 ? 23 : (lane < 6 ? 29 : 31));   // the zero entries
        if (lane < 8) rw[2 * zc + (lane & 1)] = 0.f;
    }

    // this lane's operand roles: L row a = q (its x powers and middle-row
    // offset), N column b = q (its y powers and middle-column offset)
    const float xa = static_cast<float>(min(q, P.vy - 1) - P.my) * (1.f / 16.f);
    const float yb = static_cast<float>(min(q, P.vz - 1) - P.mz) * (1.f / 16.f);
    // The product's 16 k slots: half g of the wave holds four powers m of
    // the patch polynomial, m in {0, 3, 5, 6} (g = 0) or {1, 2, 4} and a pad
    // (g = 1), which balances the Horner sums G_m(y) = sum_{n <= 6 - m} M_mn
    // y^n between the halves; slot j < 4 is the main term of the j-th power,
    // slot 4 + j its cross term.  L(a, k) = alpha x^m (main) or alpha x^{m+1}
    // (cross); each lane's four powers, the record offset of M_m0 and the
    // largest n of each.
    constexpr int GM_MA[4] = {0, 3, 5, 6}, GM_MB[4] = {1, 2, 4, 6};
    constexpr int GM_HL[4] = {7, 5, 3, 1};         // Horner length of slot s (the larger half's)
    int hbase[4];
    float lx[8];
    {
        float xp[8];
        xp[0] = 1.f;
#pragma unroll
        for (int m = 1; m < 8; ++m) xp[m] = xp[m - 1] * xa;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            lx[j] = g ? xp[GM_MB[j]] : xp[GM_MA[j]];
            lx[4 + j] = g ? xp[GM_MB[j] + 1] : xp[GM_MA[j] + 1];
            hbase[j] = g ? GM_PB[j == 3 ? 7 : GM_MB[j]] : GM_PB[GM_MA[j]];   // the pad: the zero entry
        }
    }
    const int pmid = (P.z0 + P.mz) * Ny + P.y0 + P.my;
    const float4 dr = A.dup[(P.z0 + P.mz) * Ny + P.y0 + min(q, P.vy - 1)];
    const float4 dc = A.dup[(P.z0 + min(q, P.vz - 1)) * Ny + P.y0 + P.my];
    // bilinear cross term from the corners, and the model's largest
    // deviation at this wave's pixels
    float hx = 0.f, hy = 0.f, hz = 0.f, dev = 0.f;
    {
        const int c00 = P.z0 * Ny + P.y0, c10 = c00 + P.vy - 1, c01 = c00 + (P.vz - 1) * Ny,
                  c11 = c01 + P.vy - 1;
        const float area = static_cast<float>(P.vy - 1) * static_cast<float>(P.vz - 1);
        if (area > 0.f) {
            const float4 d00 = A.dup[c00], d10 = A.dup[c10], d01 = A.dup[c01], d11 = A.dup[c11];
            const float ia = 1.f / area;
            hx = (d11.x - d10.x - d01.x + d00.x) * ia;
            hy = (d11.y - d10.y - d01.y + d00.y) * ia;
            hz = (d11.z - d10.z - d01.z + d00.z) * ia;
        }
        const float4 dcq = dc;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int a = min(arow(r), P.vy - 1);
            const float4 du = A.dup[pix(r)];
            const float4 dra = A.dup[(P.z0 + P.mz) * Ny + P.y0 + a];
            const float w = static_cast<float>(a - P.my) * static_cast<float>(min(q, P.vz - 1) - P.mz);
            const float ex = du.x - dra.x - dcq.x - w * hx;
            const float ey = du.y - dra.y - dcq.y - w * hy;
            const float ez = du.z - dra.z - dcq.z - w * hz;
            dev = fmaxf(dev, sqrtf(fmaf(ex, ex, fmaf(ey, ey, ez * ez))));
        }
        for (int o = 32; o > 0; o >>= 1) dev = fmaxf(dev, __shfl_xor(dev, o, 64));
    }
    // the separable form's bounds per unit, on squares: K |dO| dev <= 1e-6
    // and the cross term at the corners, (2 pi K |dO| |h| 256)^2 / 2 <= 1e-6
    const float dev2 = dev * dev;
    const float cx2 = 0.5f * (TWO_PI_F * 256.f) * (TWO_PI_F * 256.f) * fmaf(hx, hx, fmaf(hy, hy, hz * hz));
    (void)pmid;

    floatx16 sr, si;
#pragma unroll
    for (int r = 0; r < 16; ++r) { sr[r] = 0.f; si[r] = 0.f; }
    double2* out = A.partial + (static_cast<int64_t>(frame) * A.n_slices + slice) * A.npix;
    bool first = A.accumulate == 0;
    int since = 0;
    int acc_species = 0;
    float kos_x = 0.f, kos_y = 0.f, kos_z = 0.f;   // K O_s, the slice's reference origin
    auto flush = [&]() {
        const float2* __restrict__ Gm = A.gmma + static_cast<int64_t>(acc_species) * A.npix;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            if (live && arow(r) < P.vy && q < P.vz) {
                const int p = prow + arow(r);
                const float4 du = A.dup[p];
                const float ph = -fmaf(kos_z, du.z, fmaf(kos_y, du.y, kos_x * du.x));
                const float cs = __builtin_amdgcn_cosf(ph), sn = __builtin_amdgcn_sinf(ph);
                const float zr = cs * sr[r] - sn * si[r], zi = cs * si[r] + sn * sr[r];
                const float2 gg = Gm[p];
                double2* o = out + p;
                double2 v = first ? make_double2(0.0, 0.0) : *o;
                v.x += static_cast<double>(gg.x * zr - gg.y * zi);
                v.y += static_cast<double>(gg.x * zi + gg.y * zr);
                *o = v;
            }
            sr[r] = 0.f;
            si[r] = 0.f;
        }
        first = false;
        since = 0;
    };
    const float* const recs = reinterpret_cast<const float*>(A.mom);
    auto rec_at = [&](int t) {
        return recs[((static_cast<int64_t>(t - A.t_begin) * A.n_frames + frame) * A.n_patches + pc) *
                        (2 * GM_REC) + lane];
    };
    GmHeadWalk walk(A.heads, t0, t1, lane);
    int tc, gc;
    bool have = walk.next(tc, gc);
    float rec = have ? rec_at(tc) : 0.f;
    bool ref_set = false;
    while (have) {
        int tn, gn;
        const bool hn_ = walk.next(tn, gn);
        const float rec_next = hn_ ? rec_at(tn) : 0.f;     // in flight during this unit
        rw[dpos] = rec;
        gm_wave_sync();
        const int species = __float_as_int(rw[GM_EVX + 4]);
        if (since > 0 && species != acc_species) flush();
        acc_species = species;
        const float pht = rw[GM_EVX], kox = rw[GM_EVX + 1];
        const float koy = rw[GM_EVX + 2], koz = rw[GM_EVX + 3];
        if (!ref_set) {
            kos_x = kox;
            kos_y = koy;
            kos_z = koz;
            ref_set = true;
        }
        const float dox = kox - kos_x, doy = koy - kos_y, doz = koz - kos_z;
        const float kdo2 = fmaf(dox, dox, fmaf(doy, doy, doz * doz));
        const bool sep = !A.pixel_eval && kdo2 * dev2 <= 1e-12f && cx2 * kdo2 <= 1e-6f;
        if (!sep && live && lane == 0 && A.fallbacks) atomicAdd(A.fallbacks, 1u);
        if (sep) {
            // G of this lane's four powers at its column y = yb (a half's
            // terms past its own degree read the zero pads)
            float Gr[4], Gi[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                float gr = 0.f, gi = 0.f;
#pragma unroll
                for (int n = GM_HL[j] - 1; n >= 0; --n) {
                    gr = fmaf(gr, yb, rw[2 * (hbase[j] + n)]);
                    gi = fmaf(gi, yb, rw[2 * (hbase[j] + n) + 1]);
                }
                Gr[j] = gr;
                Gi[j] = gi;
            }
            const float tha = -fmaf(doz, dr.z, fmaf(doy, dr.y, dox * dr.x));
            const float ar = __builtin_amdgcn_cosf(tha), ai = __builtin_amdgcn_sinf(tha);
            const float thb = pht - fmaf(doz, dc.z, fmaf(doy, dc.y, dox * dc.x));
            const float br = __builtin_amdgcn_cosf(thb), bi = __builtin_amdgcn_sinf(thb);
            const float kap = TWO_PI_F * fmaf(doz, hz, fmaf(doy, hy, dox * hx));
            // cross factor -256 i kappa y
            const float cf = -256.f * kap * yb;
            // N(k, b): slot j < 4: T = beta G of the j-th power; slot 4 + j its
            // cross value cf i T = (-cf Ti, cf Tr)
            float tr[4], ti[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                tr[j] = br * Gr[j] - bi * Gi[j];
                ti[j] = br * Gi[j] + bi * Gr[j];
            }
            // L = alpha x^m (slots 0..7), N: slots 0..3 T, slots 4..7 cf i T
            // = (-cf Ti, cf Tr); the products split as they are formed.  Only
            // slots 0, 1 (powers 0, 3 and 1, 2) keep lo halves: slots 2, 3
            // hold powers m >= 4 (5, 6 and 4, none), whose terms are at most
            // phi^m / m! <= 2.0e-3 of the sum's scale (phi <= 0.47, the
            // degree-6 bound), and the cross slots (4..7) terms at most cmax <=
            // 1.4e-3 of the main ones (the bound above); their lo halves, 2^-11
            // of that, stay near 1e-6 and below 7e-7 of it
            uintx4 lrh, lrl, lih, lil, nrh, nrl, nih, nil;
#pragma unroll
            for (int u2 = 0; u2 < 4; ++u2) {
                unsigned int h, l;
                if (u2 == 0) {
                    gm_split_prod_pair(ar, lx[0], lx[1], h, l); lrh[0] = h; lrl[0] = l;
                    gm_split_prod_pair(ai, lx[0], lx[1], h, l); lih[0] = h; lil[0] = l;
                    gm_split_pair(tr[0], tr[1], h, l); nrh[0] = h; nrl[0] = l;
                    gm_split_pair(ti[0], ti[1], h, l); nih[0] = h; nil[0] = l;
                } else if (u2 == 1) {
                    lrh[1] = gm_prod_hi_pair(ar, lx[2], lx[3]);
                    lih[1] = gm_prod_hi_pair(ai, lx[2], lx[3]);
                    halfx2 hr, hi2;
                    hr[0] = static_cast<__fp16>(tr[2]);
                    hr[1] = static_cast<__fp16>(tr[3]);
                    hi2[0] = static_cast<__fp16>(ti[2]);
                    hi2[1] = static_cast<__fp16>(ti[3]);
                    nrh[1] = __builtin_bit_cast(unsigned int, hr);
                    nih[1] = __builtin_bit_cast(unsigned int, hi2);
                    lrl[1] = 0u;
                    lil[1] = 0u;
                    nrl[1] = 0u;
                    nil[1] = 0u;
                } else {
                    const int j = 2 * (u2 - 2);
                    lrh[u2] = gm_prod_hi_pair(ar, lx[2 * u2], lx[2 * u2 + 1]);
                    lih[u2] = gm_prod_hi_pair(ai, lx[2 * u2], lx[2 * u2 + 1]);
                    nrh[u2] = gm_prod_hi_pair(-cf, ti[j], ti[j + 1]);
                    nih[u2] = gm_prod_hi_pair(cf, tr[j], tr[j + 1]);
                    lrl[u2] = 0u;
                    lil[u2] = 0u;
                    nrl[u2] = 0u;
                    nil[u2] = 0u;
                }
            }
            const uintx4 sgn = {0x80008000u, 0x80008000u, 0x80008000u, 0x80008000u};
            const halfx8 Lrh = __builtin_bit_cast(halfx8, lrh), Lrl = __builtin_bit_cast(halfx8, lrl);
            const halfx8 Lih = __builtin_bit_cast(halfx8, lih), Lil = __builtin_bit_cast(halfx8, lil);
            const halfx8 Mih = __builtin_bit_cast(halfx8, lih ^ sgn), Mil = __builtin_bit_cast(halfx8, lil ^ sgn);
            const halfx8 Nrh = __builtin_bit_cast(halfx8, nrh), Nrl = __builtin_bit_cast(halfx8, nrl);
            const halfx8 Nih = __builtin_bit_cast(halfx8, nih), Nil = __builtin_bit_cast(halfx8, nil);
            // Re += Lr Nr - Li Ni, Im += Lr Ni + Li Nr
            sr = __builtin_amdgcn_mfma_f32_32x32x16_f16(Lrh, Nrh, sr, 0, 0, 0);
            si = __builtin_amdgcn_mfma_f32_32x32x16_f16(Lrh, Nih, si, 0, 0, 0);
            sr = __builtin_amdgcn_mfma_f32_32x32x16_f16(Mih, Nih, sr, 0, 0, 0);
            si = __builtin_amdgcn_mfma_f32_32x32x16_f16(Lih, Nrh, si, 0, 0, 0);
            sr = __builtin_amdgcn_mfma_f32_32x32x16_f16(Lrh, Nrl, sr, 0, 0, 0);
            si = __builtin_amdgcn_mfma_f32_32x32x16_f16(Lrh, Nil, si, 0, 0, 0);
            sr = __builtin_amdgcn_mfma_f32_32x32x16_f16(Mih, Nil, sr, 0, 0, 0);
            si = __builtin_amdgcn_mfma_f32_32x32x16_f16(Lih, Nrl, si, 0, 0, 0);
            sr = __builtin_amdgcn_mfma_f32_32x32x16_f16(Lrl, Nrh, sr, 0, 0, 0);
            si = __builtin_amdgcn_mfma_f32_32x32x16_f16(Lrl, Nih, si, 0, 0, 0);
            sr = __builtin_amdgcn_mfma_f32_32x32x16_f16(Mil, Nih, sr, 0, 0, 0);
            si = __builtin_amdgcn_mfma_f32_32x32x16_f16(Lil, Nrh, si, 0, 0, 0);
        } else {
            // exact phase per pixel: G at this lane's column is G_m(yb) (the
            // lane's pixels share its column), Horner's rule in x per pixel
            float Gr[D + 1], Gi[D + 1];
#pragma unroll
            for (int m = 0; m <= D; ++m) {
                float gr = 0.f, gi = 0.f;
#pragma unroll
                for (int n = D - m; n >= 0; --n) {
                    gr = fmaf(gr, yb, rw[2 * (GM_PB[m] + n)]);
                    gi = fmaf(gi, yb, rw[2 * (GM_PB[m] + n) + 1]);
                }
                Gr[m] = gr;
                Gi[m] = gi;
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float x = static_cast<float>(arow(r) - P.my) * (1.f / 16.f);
                float pr = Gr[D], pi = Gi[D];
#pragma unroll
                for (int m = D - 1; m >= 0; --m) {
                    pr = fmaf(pr, x, Gr[m]);
                    pi = fmaf(pi, x, Gi[m]);
                }
                const float4 du = A.dup[pix(r)];
                const float phs = pht - fmaf(doz, du.z, fmaf(doy, du.y, dox * du.x));
                const float sn = __builtin_amdgcn_sinf(phs);
                const float cs = __builtin_amdgcn_cosf(phs);
                sr[r] = fmaf(cs, pr, fmaf(-sn, pi, sr[r]));
                si[r] = fmaf(cs, pi, fmaf(sn, pr, si[r]));
            }
        }
        gm_wave_sync();                            // the record buffer's next write
        if (++since == POLY_FLUSH) flush();
        have = hn_;
        tc = tn;
        rec = rec_next;
    }
    if (since > 0) flush();
    if (live && first) {                          // no unit of this slice in the first round
#pragma unroll
        for (int r = 0; r < 16; ++r)
            if (arow(r) < P.vy && q < P.vz) out[prow + arow(r)] = make_double2(0.0, 0.0);
    }
}

// ------------------------------------------------------------------ adjoint
// Gradient of a real loss L(A) with respect to the sample-frame displacement
// of every atom, given g = dL/dRe(A) + i dL/dIm(A) per frame and pixel.  With
// c_j the atom's term of A and phi its phase,
//     dL/du_j = sum_{f,p} Re(conj(g) i c_j) dphi/du_j
//             = -2 pi K' Im(w_j Z_j),   Z_j = sum_{f,p} v B e^{2 pi i v.sigma_j},
//     B = conj(g) G e^{2 pi i Theta},  K' = 1e-10 / lambda (per Angstrom),
// where v = R^T (x_hat - u_hat) is also the phase gradient direction.  One
// workgroup owns one tile (its 256 atoms arrive as wave-uniform scalar loads);
// lanes walk the detector in KP-pixel row blocks with the same line model as
// the forward recurrence, so per atom a lane forms S0 = sum_m B_m e_m and
// S1 = sum_m m B_m e_m and Z = v0 S0 + g S1.  The 6 floats of Z are summed over
// the wave with DPP row shifts/broadcasts and accumulated per wave in FP64 in
// LDS; the tile's atoms are written once at the end (no atomics, fixed order).
struct AdjArgs {
    const AtomRec* atoms;
    const TileRec* tiles;
    const FrameRec* frames;
    const double4* u0h;
    const double* theta0;
    const float2* gtab;
    const float2* gin;     // [frame][pixel] dL/dA in the Re + i Im convention
    const float* amax;     // [2][AMAX_BLOCKS] block maxima of |gin| and |gtab|
    float* grad;           // [atom][3], accumulated into
    int n_frames;
    int Ny;
    int Nz;
    int npix;
    int nbx;
    double K;
};

// Sum over the 64 lanes of a wave; the total lands in lane 63.
__device__ __forceinline__ float wave_sum_to_63(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x112, 0xf, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x114, 0xf, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x118, 0xf, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x142, 0xa, 0xf, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x143, 0xc, 0xf, false));
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Block maxima of |x| over the components of dL/dA and of the per-species
// factors G.  Their product is ~1e-40 for fields in r_e units, below FP32's
// normal range, so the adjoint scales both to O(1) by powers of two.
__global__ __launch_bounds__(256) void k_amax(const float* __restrict__ gin, int64_t n_gin,
                                              const float* __restrict__ gtab, int64_t n_gtab,
                                              float* __restrict__ out) {
    __shared__ double lds[4];
    float a = 0.f, b = 0.f;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_gin; i += stride)
        a = fmaxf(a, fabsf(gin[i]));
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n_gtab; i += stride)
        b = fmaxf(b, fabsf(gtab[i]));
    const double ma = block_max(static_cast<double>(a), lds);
    const double mb = block_max(static_cast<double>(b), lds);
    if (threadIdx.x == 0) {
        out[blockIdx.x] = static_cast<float>(ma);
        out[AMAX_BLOCKS + blockIdx.x] = static_cast<float>(mb);
    }
}

// 2^-e with x 2^-e in [0.5, 1); 1 for x = 0 (or non-finite x).
__device__ __forceinline__ float pow2_scale(float x) {
    if (!(x > 0.f) || !isfinite(x)) return 1.f;
    int e;
    frexpf(x, &e);
    return ldexpf(1.f, -e);
}

template <int KP, bool SECOND>
__global__ __launch_bounds__(SUM_THREADS) void k_adjoint(const AdjArgs A) {
    __shared__ double acc[SUM_THREADS / 64][TILE][6];
    __shared__ double red[SUM_THREADS / 64];
    const int t = blockIdx.x;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int i = tid; i < (SUM_THREADS / 64) * TILE * 6; i += SUM_THREADS)
        (&acc[0][0][0])[i] = 0.0;
    float mg = 0.f, mG = 0.f;
    for (int i = tid; i < AMAX_BLOCKS; i += SUM_THREADS) {
        mg = fmaxf(mg, A.amax[i]);
        mG = fmaxf(mG, A.amax[AMAX_BLOCKS + i]);
    }
    const float sg = pow2_scale(static_cast<float>(block_max(static_cast<double>(mg), red)));
    const float sG = pow2_scale(static_cast<float>(block_max(static_cast<double>(mG), red)));
    __syncthreads();
    const TileRec tr = A.tiles[t];
    const AtomRec* __restrict__ at = A.atoms + static_cast<int64_t>(t) * TILE;
    const int cnt = tr.count;
    const float2* __restrict__ G = A.gtab + static_cast<int64_t>(tr.species) * A.npix;
    const int lanes = A.Nz * A.nbx;
    const int passes = (lanes + SUM_THREADS - 1) / SUM_THREADS;
    constexpr float TWO_PI_F = 6.2831853071795865f;

    for (int f = 0; f < A.n_frames; ++f) {
        const FrameRec& F = A.frames[f];
        const double R00 = F.R[0], R01 = F.R[1], R02 = F.R[2];
        const double R10 = F.R[3], R11 = F.R[4], R12 = F.R[5];
        const double R20 = F.R[6], R21 = F.R[7], R22 = F.R[8];
        const double Ox = fma(R00, tr.ox, fma(R01, tr.oy, fma(R02, tr.oz, F.T[0])));
        const double Oy = fma(R10, tr.ox, fma(R11, tr.oy, fma(R12, tr.oz, F.T[1])));
        const double Oz = fma(R20, tr.ox, fma(R21, tr.oy, fma(R22, tr.oz, F.T[2])));
        const double O2 = Ox * Ox + Oy * Oy + Oz * Oz;
        const float xs0 = static_cast<float>(F.xs[0]);
        const float xs1 = static_cast<float>(F.xs[1]);
        const float xs2 = static_cast<float>(F.xs[2]);
        const float2* __restrict__ gin = A.gin + static_cast<int64_t>(f) * A.npix;
        for (int pass = 0; pass < passes; ++pass) {
            const int lb = pass * SUM_THREADS + tid;
            const bool lane_ok = lb < lanes;
            const int lbc = lane_ok ? lb : 0;
            const int iz = lbc / A.nbx;
            const int iy0 = (lbc % A.nbx) * KP;
            const int nv = min(KP, A.Ny - iy0);
            const int pbase = iz * A.Ny + iy0;
            const int m_last = nv - 1, m_a = (nv - 1) / 2, m_b = nv / 2;
            auto vdir = [&](int m, double& sx, double& sy, double& sz) {
                const double4 u = A.u0h[pbase + m];
                const double uo = u.x * Ox + u.y * Oy + u.z * Oz;
                const double h2 = 2.0 * u.w;
                const double wx = 1.0 - u.x + (Ox - uo * u.x) * h2;
                const double wy = -u.y + (Oy - uo * u.y) * h2;
                const double wz = -u.z + (Oz - uo * u.z) * h2;
                sx = R00 * wx + R10 * wy + R20 * wz;
                sy = R01 * wx + R11 * wy + R21 * wz;
                sz = R02 * wx + R12 * wy + R22 * wz;
            };
            double a0x, a0y, a0z, a1x, a1y, a1z, max_, may_, maz_, mbx_, mby_, mbz_;
            vdir(0, a0x, a0y, a0z);
            vdir(m_last, a1x, a1y, a1z);
            vdir(m_a, max_, may_, maz_);
            vdir(m_b, mbx_, mby_, mbz_);
            const double inv_n = m_last > 0 ? 1.0 / m_last : 0.0;
            const double gapx = 0.5 * (a0x + a1x) - 0.5 * (max_ + mbx_);
            const double gapy = 0.5 * (a0y + a1y) - 0.5 * (may_ + mby_);
            const double gapz = 0.5 * (a0z + a1z) - 0.5 * (maz_ + mbz_);
            const float v0x = static_cast<float>(a0x - 0.5 * gapx);
            const float v0y = static_cast<float>(a0y - 0.5 * gapy);
            const float v0z = static_cast<float>(a0z - 0.5 * gapz);
            const float gx = static_cast<float>((a1x - a0x) * inv_n);
            const float gy = static_cast<float>((a1y - a0y) * inv_n);
            const float gz = static_cast<float>((a1z - a0z) * inv_n);
            const float hh = SECOND ? static_cast<float>(A.u0h[pbase + m_a].w / A.K) : 0.f;
            const float hmid = 0.5f * static_cast<float>(m_last);
            // B_m = conj(g) G e^{2 pi i Theta} (scaled by sg sG); zero for
            // pixels past the row end
            float br[KP], bi[KP];
#pragma unroll
            for (int m = 0; m < KP; ++m) {
                const int p = pbase + min(m, m_last);
                const double4 u = A.u0h[p];
                const double uo = u.x * Ox + u.y * Oy + u.z * Oz;
                double th = A.theta0[p] + A.K * (Ox - uo + (O2 - uo * uo) * u.w);
                th -= rint(th);
                const float thf = __builtin_amdgcn_fractf(static_cast<float>(th));
                const float sn = __builtin_amdgcn_sinf(thf);
                const float cs = __builtin_amdgcn_cosf(thf);
                const float2 g0 = G[p];
                const float2 gg = make_float2(g0.x * sG, g0.y * sG);
                const float zr = gg.x * cs - gg.y * sn, zi = gg.x * sn + gg.y * cs;
                const float2 q0 = gin[p];                      // conj(q) * z
                const float2 q = make_float2(q0.x * sg, q0.y * sg);
                const bool live = lane_ok && m < nv;
                br[m] = live ? q.x * zr + q.y * zi : 0.f;
                bi[m] = live ? q.x * zi - q.y * zr : 0.f;
            }
            AtomRec nxt = at[0];
            for (int j = 0; j < cnt; ++j) {
                const AtomRec a = nxt;
                if (j + 1 < cnt) nxt = at[j + 1];
                float psi = fmaf(v0z, a.sz, fmaf(v0y, a.sy, v0x * a.sx));
                const float d = fmaf(gz, a.sz, fmaf(gy, a.sy, gx * a.sx));
                if (SECOND) {
                    const float u =
                        fmaf(xs2, a.sz, fmaf(xs1, a.sy, xs0 * a.sx)) - fmaf(hmid, d, psi);
                    psi = fmaf(hh, fmaf(-u, u, a.s2), psi);
                }
                const float r = __builtin_amdgcn_fractf(psi);
                float er = __builtin_amdgcn_cosf(r);
                float ei = __builtin_amdgcn_sinf(r);
                const float x = d * TWO_PI_F;
                const float x2 = x * x;
                const float ar = fmaf(x2, fmaf(x2, 0.041666668f, -0.5f), 1.0f);
                const float ai = x * fmaf(x2, -0.16666667f, 1.0f);
                float s0r = 0.f, s0i = 0.f, s1r = 0.f, s1i = 0.f;
#pragma unroll
                for (int m = 0; m < KP; ++m) {
                    const float tr_ = fmaf(br[m], er, -bi[m] * ei);
                    const float ti_ = fmaf(br[m], ei, bi[m] * er);
                    s0r += tr_;
                    s0i += ti_;
                    if (m > 0) {
                        s1r = fmaf(static_cast<float>(m), tr_, s1r);
                        s1i = fmaf(static_cast<float>(m), ti_, s1i);
                    }
                    if (m + 1 < KP) {
                        const float nr = fmaf(er, ar, -ei * ai);
                        ei = fmaf(er, ai, ei * ar);
                        er = nr;
                    }
                }
                float z[6];
                z[0] = fmaf(v0x, s0r, gx * s1r);
                z[1] = fmaf(v0x, s0i, gx * s1i);
                z[2] = fmaf(v0y, s0r, gy * s1r);
                z[3] = fmaf(v0y, s0i, gy * s1i);
                z[4] = fmaf(v0z, s0r, gz * s1r);
                z[5] = fmaf(v0z, s0i, gz * s1i);
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    const float tot = wave_sum_to_63(z[c]);
                    if (lane == 0) acc[wave][j][c] += static_cast<double>(tot);
                }
            }
        }
    }
    __syncthreads();
    if (tid < cnt) {
        const AtomRec a = at[tid];
        double Z[6];
        for (int c = 0; c < 6; ++c) {
            Z[c] = 0.0;
            for (int w = 0; w < SUM_THREADS / 64; ++w) Z[c] += acc[w][tid][c];
        }
        const double scale = -2.0 * PI * A.K * 1e-10 /
                             (static_cast<double>(sg) * static_cast<double>(sG));
        const uint32_t idx = __float_as_uint(a.pad0);
        double wi, wr;
        sincospi(2.0 * static_cast<double>(a.ph), &wi, &wr);
        wr *= a.mag;
        wi *= a.mag;
        for (int c = 0; c < 3; ++c) {
            const double g = scale * (wr * Z[2 * c + 1] + wi * Z[2 * c]);
            A.grad[3 * static_cast<int64_t>(idx) + c] += static_cast<float>(g);
        }
    }
}

// Matrix-core adjoint.  With the forward's separable patch model the atom
// phasor is A_j(a) Bz_j(b), A_j(a) = e^{2 pi i va(a).sigma_j} (+ the atom's
// second-order phase), Bz_j(b) = e^{2 pi i wb(b).sigma_j}, and the gradient
// direction is v(a, b) = va(a) + wb(b).  Per frame, patch and tile, with the
// pixel weights W(a, b) = conj(g) G' e^{2 pi i (tile origin phase)},
//     Z_j = sum_ab v(a, b) W(a, b) A_j(a) Bz_j(b)
//         = vc S0 + gy S1 + hy S2 + gz U1 + hz U2,
//     S_n = sum_a a^n A_j(a) Y0_j(a),  U_n = sum_a A_j(a) Y_n_j(a),
//     Y_n_j(a) = sum_b W(a, b) b^n Bz_j(b)   (a, b offsets from the middle),
// so NV complex products per block of 32 atoms, W (32 x 32, the same for
// every atom) times [Bz, b Bz, b^2 Bz] (32 x 32 atoms), on the FP16-split
// matrix cores as in the forward, followed by a contraction over a in the
// lanes that hold each atom's column.  One workgroup owns one tile; its four
// waves split the (frame, patch) pairs and accumulate Z per atom in FP64 in
// LDS; the tile's atoms are written once at the end.  NV = 2 leaves out
// hz U2 (the host takes it where |hz| b^2 <= 1e-6 over the patch, a column
// curvature that fine-pitch detectors keep far below that); NV = 3 keeps it.
struct AdjMmaArgs {
    const AtomRec* atoms;
    const TileRec* tiles;
    const FrameRec* frames;
    const double4* u0h;
    const double* theta0;
    const float4* dup;
    const float2* gmma;
    const float2* gin;     // [frame][pixel] dL/dA in the Re + i Im convention
    const float* amax;     // [2][AMAX_BLOCKS] block maxima of |gin| and |G|
    float* grad;           // [atom][3], accumulated into
    int n_frames;
    int Ny;
    int Nz;
    int npix;
    int pby;
    int n_patches;
    double K;
    // moment adjoint over units of tiles (k_adjoint_moments, k_adjoint_poly<..., true>)
    const float* grp_r;    // staged pair / quad radii ([tile][2])
    float grp_rmax;        // largest unit radius inside the error bounds
    int n_tiles;
    float* mom;            // [tile][frame x patch][2 NM] pixel moments of unit heads
};

template <bool SECOND, int NV>
__global__ __launch_bounds__(64 * MMA_WAVES) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_adjoint_mma(const AdjMmaArgs A) {
    __shared__ AtomRec tile[TILE];
    __shared__ double zacc[MMA_WAVES][TILE][6];
    __shared__ double red[MMA_WAVES];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int t = blockIdx.x;
    tile[tid] = A.atoms[static_cast<int64_t>(t) * TILE + tid];
    for (int i = tid; i < MMA_WAVES * TILE * 6; i += 64 * MMA_WAVES) (&zacc[0][0][0])[i] = 0.0;
    float mg = 0.f, mG = 0.f;
    for (int i = tid; i < AMAX_BLOCKS; i += 64 * MMA_WAVES) {
        mg = fmaxf(mg, A.amax[i]);
        mG = fmaxf(mG, A.amax[AMAX_BLOCKS + i]);
    }
    const float sg = pow2_scale(static_cast<float>(block_max(static_cast<double>(mg), red)));
    const float sG = pow2_scale(static_cast<float>(block_max(static_cast<double>(mG), red)));
    __syncthreads();

    const TileRec tr = A.tiles[t];
    const int cnt = tr.count;
    const int nblk = (cnt + 31) >> 5;
    const float2* __restrict__ Gt = A.gmma + static_cast<int64_t>(tr.species) * A.npix;
    const int h = lane >> 5, q = lane & 31;
    const int Ny = A.Ny;
    const int npairs = A.n_frames * A.n_patches;

    for (int pr = wave; pr < npairs; pr += MMA_WAVES) {
        const int f = pr / A.n_patches, patch = pr % A.n_patches;
        const FrameRec& F = A.frames[f];
        const PatchGeom P = patch_of(patch % A.pby, patch / A.pby, A.Ny, A.Nz);
        const double R00 = F.R[0], R01 = F.R[1], R02 = F.R[2];
        const double R10 = F.R[3], R11 = F.R[4], R12 = F.R[5];
        const double R20 = F.R[6], R21 = F.R[7], R22 = F.R[8];
        const int pmid = (P.z0 + P.mz) * Ny + P.y0 + P.my;

        // direction model (sample frame) from the exact direction at the
        // middle pixel and the ends of the middle row and column
        float vc[3], gy[3], hy[3], gz[3], hz[3];
        double vc0[3];
        {
            const int samp[5] = {pmid, pmid - P.my, pmid - P.my + P.vy - 1, pmid - P.mz * Ny,
                                 pmid + (P.vz - 1 - P.mz) * Ny};
            double vs[5][3];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const double4 u = A.u0h[samp[k]];
                const double wx = 1.0 - u.x, wy = -u.y, wz = -u.z;
                vs[k][0] = R00 * wx + R10 * wy + R20 * wz;
                vs[k][1] = R01 * wx + R11 * wy + R21 * wz;
                vs[k][2] = R02 * wx + R12 * wy + R22 * wz;
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double gq, hq;
                quad_fit(vs[1][d] - vs[0][d], vs[2][d] - vs[0][d], -P.my, P.vy - 1 - P.my, gq, hq);
                gy[d] = static_cast<float>(gq);
                hy[d] = static_cast<float>(hq);
                quad_fit(vs[3][d] - vs[0][d], vs[4][d] - vs[0][d], -P.mz, P.vz - 1 - P.mz, gq, hq);
                gz[d] = static_cast<float>(gq);
                hz[d] = static_cast<float>(hq);
                vc0[d] = vs[0][d];
            }
        }
        // this tile's origin: phase at the middle pixel, direction shift
        const double4 um = A.u0h[pmid];
        const double Ox = fma(R00, tr.ox, fma(R01, tr.oy, fma(R02, tr.oz, F.T[0])));
        const double Oy = fma(R10, tr.ox, fma(R11, tr.oy, fma(R12, tr.oz, F.T[1])));
        const double Oz = fma(R20, tr.ox, fma(R21, tr.oy, fma(R22, tr.oz, F.T[2])));
        const double O2 = Ox * Ox + Oy * Oy + Oz * Oz;
        const double uo = um.x * Ox + um.y * Oy + um.z * Oz;
        double ph_t = A.theta0[pmid] + A.K * (Ox - uo + (O2 - uo * uo) * um.w);
        ph_t -= floor(ph_t);
        const double h2 = 2.0 * um.w;
        const double cxl = (Ox - uo * um.x) * h2, cyl = (Oy - uo * um.y) * h2,
                     czl = (Oz - uo * um.z) * h2;
        vc[0] = static_cast<float>(vc0[0] + (R00 * cxl + R10 * cyl + R20 * czl));
        vc[1] = static_cast<float>(vc0[1] + (R01 * cxl + R11 * cyl + R21 * czl));
        vc[2] = static_cast<float>(vc0[2] + (R02 * cxl + R12 * cyl + R22 * czl));
        const float xc0 = static_cast<float>(F.xs[0] - vc0[0]);
        const float xc1 = static_cast<float>(F.xs[1] - vc0[1]);
        const float xc2 = static_cast<float>(F.xs[2] - vc0[2]);
        const float hh = SECOND ? static_cast<float>(um.w / A.K) : 0.f;
        const float kox = static_cast<float>(A.K * Ox), koy = static_cast<float>(A.K * Oy),
                    koz = static_cast<float>(A.K * Oz);
        const float pht = static_cast<float>(ph_t);

        // W(a = q, b) for this lane's b = 16 s + 8 h + u, split into FP16
        // halves (zero outside the detector)
        halfx8 wrh[2], wrl[2], wih[2], wil[2];
        const float2* __restrict__ gf = A.gin + static_cast<int64_t>(f) * A.npix;
#pragma unroll
        for (int st = 0; st < 2; ++st) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if ((u & 3) == 0) __builtin_amdgcn_sched_barrier(0);   // bounds loads in flight
                const int b = 16 * st + 8 * h + u;
                const bool okp = q < P.vy && b < P.vz;
                const int p = (P.z0 + min(b, P.vz - 1)) * Ny + P.y0 + min(q, P.vy - 1);
                float wr = 0.f, wi = 0.f;
                if (okp) {
                    const float4 du = A.dup[p];
                    const float phs = __builtin_amdgcn_fractf(
                        pht - fmaf(koz, du.z, fmaf(koy, du.y, kox * du.x)));
                    const float sn = __builtin_amdgcn_sinf(phs), cs = __builtin_amdgcn_cosf(phs);
                    const float2 g0 = Gt[p], q0 = gf[p];
                    const float gr = g0.x * sG, gi = g0.y * sG, qr = q0.x * sg, qi = q0.y * sg;
                    const float zr = gr * cs - gi * sn, zi = gr * sn + gi * cs;
                    wr = qr * zr + qi * zi;          // conj(g) z
                    wi = qr * zi - qi * zr;
                }
                split16(wr, wrh[st], wrl[st], u);
                split16(wi, wih[st], wil[st], u);
            }
        }

        for (int blk = 0; blk < nblk; ++blk) {
            // B operand lanes: atom blk * 32 + q (records past the count are
            // zero and their Z is not kept)
            // per-lane row and column offsets are rebuilt per block (tied to
            // blk, which is < 65536): hoisted, they would hold ~50 registers
            const int tie = blk >> 16;
            const AtomRec a = tile[blk * 32 + q];
            const float d1 = fmaf(gz[2], a.sz, fmaf(gz[1], a.sy, gz[0] * a.sx));
            const float d2 = fmaf(hz[2], a.sz, fmaf(hz[1], a.sy, hz[0] * a.sx));
            float bzr[2][8], bzi[2][8];
#pragma unroll
            for (int st = 0; st < 2; ++st)
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const float bo = static_cast<float>(16 * st + 8 * h + u - P.mz + tie);
                    const float fb = __builtin_amdgcn_fractf(bo * fmaf(bo, d2, d1));
                    bzr[st][u] = __builtin_amdgcn_cosf(fb);
                    bzi[st][u] = __builtin_amdgcn_sinf(fb);
                }
            // A_j(a) at this lane's result rows a(r) = 8 (r / 4) + 4 h + r % 4
            const float e0 = fmaf(vc[2], a.sz, fmaf(vc[1], a.sy, vc[0] * a.sx));
            const float e1 = fmaf(gy[2], a.sz, fmaf(gy[1], a.sy, gy[0] * a.sx));
            const float e2 = fmaf(hy[2], a.sz, fmaf(hy[1], a.sy, hy[0] * a.sx));
            float e2nd = 0.f;
            if (SECOND) {
                const float w = fmaf(xc2, a.sz, fmaf(xc1, a.sy, xc0 * a.sx)) - e0;
                e2nd = hh * fmaf(-w, w, a.s2);
            }
            float ar[16], ai[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float ao = static_cast<float>(8 * (r >> 2) + 4 * h + (r & 3) - P.my + tie);
                const float fa = __builtin_amdgcn_fractf(fmaf(ao, fmaf(ao, e2, e1), e0 + e2nd));
                ar[r] = __builtin_amdgcn_cosf(fa);
                ai[r] = __builtin_amdgcn_sinf(fa);
            }
            float s0r = 0.f, s0i = 0.f, s1r = 0.f, s1i = 0.f, s2r = 0.f, s2i = 0.f;
            float u1r = 0.f, u1i = 0.f, u2r = 0.f, u2i = 0.f;
#pragma unroll 1
            for (int n = 0; n < NV; ++n) {
                floatx16 crr, cii, cri, cir;
#pragma unroll
                for (int r = 0; r < 16; ++r) { crr[r] = 0.f; cii[r] = 0.f; cri[r] = 0.f; cir[r] = 0.f; }
#pragma unroll
                for (int st = 0; st < 2; ++st) {
                    halfx8 brh, brl, bih, bil;
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const float bo = static_cast<float>(16 * st + 8 * h + u - P.mz + tie + n * tie);
                        const float m = n == 0 ? 1.f : (n == 1 ? bo : bo * bo);
                        split16(m * bzr[st][u], brh, brl, u);
                        split16(m * bzi[st][u], bih, bil, u);
                    }
                    crr = __builtin_amdgcn_mfma_f32_32x32x16_f16(wrh[st], brh, crr, 0, 0, 0);
                    cii = __builtin_amdgcn_mfma_f32_32x32x16_f16(wih[st], bih, cii, 0, 0, 0);
                    cri = __builtin_amdgcn_mfma_f32_32x32x16_f16(wrh[st], bih, cri, 0, 0, 0);
                    cir = __builtin_amdgcn_mfma_f32_32x32x16_f16(wih[st], brh, cir, 0, 0, 0);
                    crr = __builtin_amdgcn_mfma_f32_32x32x16_f16(wrh[st], brl, crr, 0, 0, 0);
                    cii = __builtin_amdgcn_mfma_f32_32x32x16_f16(wih[st], bil, cii, 0, 0, 0);
                    cri = __builtin_amdgcn_mfma_f32_32x32x16_f16(wrh[st], bil, cri, 0, 0, 0);
                    cir = __builtin_amdgcn_mfma_f32_32x32x16_f16(wih[st], brl, cir, 0, 0, 0);
                    crr = __builtin_amdgcn_mfma_f32_32x32x16_f16(wrl[st], brh, crr, 0, 0, 0);
                    cii = __builtin_amdgcn_mfma_f32_32x32x16_f16(wil[st], bih, cii, 0, 0, 0);
                    cri = __builtin_amdgcn_mfma_f32_32x32x16_f16(wrl[st], bih, cri, 0, 0, 0);
                    cir = __builtin_amdgcn_mfma_f32_32x32x16_f16(wil[st], brh, cir, 0, 0, 0);
                }
                // Y_n(a(r)) = (C_rr - C_ii) + i (C_ri + C_ir); P = A Y_n
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const float yr = crr[r] - cii[r], yi = cri[r] + cir[r];
                    const float pr_ = ar[r] * yr - ai[r] * yi, pi_ = ar[r] * yi + ai[r] * yr;
                    if (n == 0) {
                        const float ao =
                            static_cast<float>(8 * (r >> 2) + 4 * h + (r & 3) - P.my + tie);
                        s0r += pr_;
                        s0i += pi_;
                        s1r = fmaf(ao, pr_, s1r);
                        s1i = fmaf(ao, pi_, s1i);
                        s2r = fmaf(ao * ao, pr_, s2r);
                        s2i = fmaf(ao * ao, pi_, s2i);
                    } else if (n == 1) {
                        u1r += pr_;
                        u1i += pi_;
                    } else {
                        u2r += pr_;
                        u2i += pi_;
                    }
                }
            }
            // the other 16 rows of this atom's column are in lane ^ 32
            s0r += __shfl_xor(s0r, 32, 64); s0i += __shfl_xor(s0i, 32, 64);
            s1r += __shfl_xor(s1r, 32, 64); s1i += __shfl_xor(s1i, 32, 64);
            s2r += __shfl_xor(s2r, 32, 64); s2i += __shfl_xor(s2i, 32, 64);
            u1r += __shfl_xor(u1r, 32, 64); u1i += __shfl_xor(u1i, 32, 64);
            if (NV > 2) {
                u2r += __shfl_xor(u2r, 32, 64); u2i += __shfl_xor(u2i, 32, 64);
            }
            const int j = blk * 32 + q;
            if (h == 0 && j < cnt) {
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    float zr = vc[d] * s0r + gy[d] * s1r + hy[d] * s2r + gz[d] * u1r;
                    float zi = vc[d] * s0i + gy[d] * s1i + hy[d] * s2i + gz[d] * u1i;
                    if (NV > 2) {
                        zr = fmaf(hz[d], u2r, zr);
                        zi = fmaf(hz[d], u2i, zi);
                    }
                    zacc[wave][j][2 * d] += static_cast<double>(zr);
                    zacc[wave][j][2 * d + 1] += static_cast<double>(zi);
                }
            }
        }
    }
    __syncthreads();
    if (tid < cnt) {
        const AtomRec a = tile[tid];
        double Z[6];
        for (int c = 0; c < 6; ++c) {
            Z[c] = 0.0;
            for (int w = 0; w < MMA_WAVES; ++w) Z[c] += zacc[w][tid][c];
        }
        const double scale = -2.0 * PI * A.K * 1e-10 /
                             (static_cast<double>(sg) * static_cast<double>(sG));
        const uint32_t idx = __float_as_uint(a.pad0);
        double wi, wr;
        sincospi(2.0 * static_cast<double>(a.ph), &wi, &wr);
        wr *= a.mag;
        wi *= a.mag;
        for (int c = 0; c < 3; ++c) {
            const double g = scale * (wr * Z[2 * c + 1] + wi * Z[2 * c]);
            A.grad[3 * static_cast<int64_t>(idx) + c] += static_cast<float>(g);
        }
    }
}

// Pixel moments of the moment adjoint over units of tiles: per unit head and
// (frame, patch) pair, M_mn = sum_ab W(a, b) x^m y^n (m + n <= D + 2), W the
// weighted residual times the unit's origin phase, formed as k_adjoint_poly
// forms them per tile; the unit's tiles then share them.
constexpr int ADJ_GROUP = 64;     // (frame, patch) pairs per k_adjoint_moments workgroup
// Pixel moments M_mn (m + n <= D + 2) per unit head and (frame, patch) pair
// at expansion degree D; the scratch is sized for the largest degree the
// host chooses (poly_degree returns 5 or 6).
constexpr int adj_moments(int D) { return (D + 3) * (D + 4) / 2; }
constexpr int ADJ_MAX_DEGREE = 6;

template <int D>
__global__ __launch_bounds__(64 * POLY_WAVES) void k_adjoint_moments(const AdjMmaArgs A) {
    static_assert(D <= ADJ_MAX_DEGREE, "the moment scratch is sized for ADJ_MAX_DEGREE");
    constexpr int DM = D + 2;
    constexpr int NM = adj_moments(D);              // pixel moments
    constexpr int ROUNDS = 3;                       // reduced in three rounds: a
    constexpr int HALF = (NM + ROUNDS - 1) / ROUNDS; // 32 KB buffer fits 5 workgroups per CU
    constexpr int NVP = (2 * HALF) | 1;
    __shared__ float red[POLY_WAVES][64 * NVP];
    __shared__ double rmx[POLY_WAVES];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int t = blockIdx.x;
    int head, size;
    unit_of(t, A.n_tiles, A.grp_r, A.grp_rmax, head, size);
    if (head != t) return;                           // the whole workgroup
    float mg = 0.f, mG = 0.f;
    for (int i = tid; i < AMAX_BLOCKS; i += 64 * POLY_WAVES) {
        mg = fmaxf(mg, A.amax[i]);
        mG = fmaxf(mG, A.amax[AMAX_BLOCKS + i]);
    }
    const float sg = pow2_scale(static_cast<float>(block_max(static_cast<double>(mg), rmx)));
    const float sG = pow2_scale(static_cast<float>(block_max(static_cast<double>(mG), rmx)));
    const TileRec tr = A.tiles[t];
    double uox = tr.ox, uoy = tr.oy, uoz = tr.oz;
    if (size > 1) {
        const TileRec b1 = A.tiles[t + 1];
        const TileRec b2 = A.tiles[size == 4 ? t + 2 : t];
        const TileRec b3 = A.tiles[size == 4 ? t + 3 : t];
        uox = unit_origin(size, tr.ox, b1.ox, b2.ox, b3.ox);
        uoy = unit_origin(size, tr.oy, b1.oy, b2.oy, b3.oy);
        uoz = unit_origin(size, tr.oz, b1.oz, b2.oz, b3.oz);
    }
    const float2* __restrict__ Gt = A.gmma + static_cast<int64_t>(tr.species) * A.npix;
    const int h = lane >> 5, q = lane & 31;
    const int Ny = A.Ny;
    const int npairs = A.n_frames * A.n_patches;
    const int p0 = blockIdx.y * ADJ_GROUP, p1 = min(npairs, p0 + ADJ_GROUP);
    float* rw = red[wave];
    // every wave runs the same number of pair iterations (barriers inside)
    const int iters = (p1 - p0 + POLY_WAVES - 1) / POLY_WAVES;
    for (int it = 0; it < iters; ++it) {
        const int pr = p0 + it * POLY_WAVES + wave;
        const bool live = pr < p1;
        const int prc = live ? pr : p0;
        const int f = prc / A.n_patches, patch = prc % A.n_patches;
        const FrameRec& F = A.frames[f];
        const PatchGeom P = patch_of(patch % A.pby, patch / A.pby, A.Ny, A.Nz);
        const int pmid = (P.z0 + P.mz) * Ny + P.y0 + P.my;
        const double4 um = A.u0h[pmid];
        const double Ox = fma(F.R[0], uox, fma(F.R[1], uoy, fma(F.R[2], uoz, F.T[0])));
        const double Oy = fma(F.R[3], uox, fma(F.R[4], uoy, fma(F.R[5], uoz, F.T[1])));
        const double Oz = fma(F.R[6], uox, fma(F.R[7], uoy, fma(F.R[8], uoz, F.T[2])));
        const double O2 = Ox * Ox + Oy * Oy + Oz * Oz;
        const double uo = um.x * Ox + um.y * Oy + um.z * Oz;
        double ph_t = A.theta0[pmid] + A.K * (Ox - uo + (O2 - uo * uo) * um.w);
        ph_t -= floor(ph_t);
        const float kox = static_cast<float>(A.K * Ox), koy = static_cast<float>(A.K * Oy),
                    koz = static_cast<float>(A.K * Oz);
        const float pht = static_cast<float>(ph_t);

        // W on this lane's row a = q, columns b = 16 s + 8 h + u, folded into
        // H_n = sum_b W y_b^n
        float Hr[DM + 1], Hi[DM + 1];
#pragma unroll
        for (int n = 0; n <= DM; ++n) { Hr[n] = 0.f; Hi[n] = 0.f; }
        const float2* __restrict__ gf = A.gin + static_cast<int64_t>(f) * A.npix;
#pragma unroll
        for (int st = 0; st < 2; ++st) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if ((u & 3) == 0) __builtin_amdgcn_sched_barrier(0);
                const int b = 16 * st + 8 * h + u;
                const bool okp = live && q < P.vy && b < P.vz;
                const int p = (P.z0 + min(b, P.vz - 1)) * Ny + P.y0 + min(q, P.vy - 1);
                float wr = 0.f, wi = 0.f;
                if (okp) {
                    const float4 du = A.dup[p];
                    const float phs = __builtin_amdgcn_fractf(
                        pht - fmaf(koz, du.z, fmaf(koy, du.y, kox * du.x)));
                    const float sn = __builtin_amdgcn_sinf(phs), cs = __builtin_amdgcn_cosf(phs);
                    const float2 g0 = Gt[p], q0 = gf[p];
                    const float gr = g0.x * sG, gi = g0.y * sG, qr = q0.x * sg, qi = q0.y * sg;
                    const float zr_ = gr * cs - gi * sn, zi_ = gr * sn + gi * cs;
                    wr = qr * zr_ + qi * zi_;          // conj(g) z
                    wi = qr * zi_ - qi * zr_;
                }
                const float y = static_cast<float>(b - P.mz) * (1.f / 16.f);
                float yn = 1.f;
#pragma unroll
                for (int n = 0; n <= DM; ++n) {
                    Hr[n] = fmaf(wr, yn, Hr[n]);
                    Hi[n] = fmaf(wi, yn, Hi[n]);
                    yn *= y;
                }
            }
        }
        // M_mn = sum over lanes of x_q^m H_n, in ROUNDS rounds through LDS; the
        // lane holding a column's sum writes it
        const float xq = static_cast<float>(q - P.my) * (1.f / 16.f);
        float xm[DM + 1];
        xm[0] = 1.f;
#pragma unroll
        for (int m = 1; m <= DM; ++m) xm[m] = xm[m - 1] * xq;
        float* out = A.mom + (static_cast<int64_t>(t) * npairs + prc) * (2 * NM);
#pragma unroll
        for (int round = 0; round < ROUNDS; ++round) {
            int k = 0;
#pragma unroll
            for (int m = 0; m <= DM; ++m) {
#pragma unroll
                for (int n = 0; n <= DM - m; ++n, ++k) {
                    if (k / HALF == round) {
                        const int c = k - round * HALF;
                        rw[lane * NVP + 2 * c] = xm[m] * Hr[n];
                        rw[lane * NVP + 2 * c + 1] = xm[m] * Hi[n];
                    }
                }
            }
            __syncthreads();
            const int col = min(lane, NVP - 1);
            float tot = 0.f;
#pragma unroll 8
            for (int l = 0; l < 64; ++l) tot += rw[l * NVP + col];
            __syncthreads();
            const int idx = round * 2 * HALF + lane;
            if (live && lane < 2 * HALF && idx < 2 * NM) out[idx] = tot;
        }
    }
}

// Adjoint by moments, the dual of k_sum_poly.  Per (frame, patch) and tile,
// with the patch's direction model v(x, y) = vc + x Gy + x^2 Hy + y Gz + y^2 Hz
// (x = a / 16, y = b / 16; Gy = 16 gy, Hy = 256 hy, ...) and the atom phase
// expansion of the moment sum, e^{2 pi i v.sigma_j} = e^{2 pi i e0_j} sum_mn c_mn(j) x^m y^n,
//     Z_j = sum_ab v(a, b) W(a, b) e^{2 pi i v.sigma_j}
//         = e^{2 pi i e0_j} sum_mn c_mn(j) (vc M_mn + Gy M_{m+1,n} + Hy M_{m+2,n}
//                                          + Gz M_{m,n+1} + Hz M_{m,n+2}),
//     M_mn = sum_ab W(a, b) x^m y^n   (m + n <= D + 2),
// so a wave forms the patch's pixel moments once (each lane the 16 pixels of
// one row a = q, summed over the wave through LDS), and every atom contracts
// its expansion coefficients with five shifted moment sets.  Z accumulates
// per lane in FP32 over up to 16 (frame, patch) pairs, then in FP64 in LDS.
template <int D, bool SECOND, bool UNITS>
__global__ __launch_bounds__(64 * POLY_WAVES) void k_adjoint_poly(const AdjMmaArgs A) {
    static_assert(D <= ADJ_MAX_DEGREE, "the moment scratch is sized for ADJ_MAX_DEGREE");
    constexpr int DM = D + 2;
    constexpr int NM = adj_moments(D);              // pixel moments
    constexpr int HALF = (NM + 1) / 2;              // reduced in two rounds
    constexpr int NVP = (2 * HALF) | 1;
    constexpr int NC = (D + 1) * (D + 2) / 2;
    constexpr float TWO_PI_F = 6.283185307179586f;
    __shared__ AtomRec tile[TILE];
    __shared__ double zacc[TILE][6];
    __shared__ float red[UNITS ? 1 : POLY_WAVES][UNITS ? 1 : 64 * NVP];
    __shared__ double rmx[POLY_WAVES];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int t = blockIdx.x;
    tile[tid] = A.atoms[static_cast<int64_t>(t) * TILE + tid];
    for (int i = tid; i < TILE * 6; i += 64 * POLY_WAVES) (&zacc[0][0])[i] = 0.0;
    float mg = 0.f, mG = 0.f;
    for (int i = tid; i < AMAX_BLOCKS; i += 64 * POLY_WAVES) {
        mg = fmaxf(mg, A.amax[i]);
        mG = fmaxf(mG, A.amax[AMAX_BLOCKS + i]);
    }
    const float sg = pow2_scale(static_cast<float>(block_max(static_cast<double>(mg), rmx)));
    const float sG = pow2_scale(static_cast<float>(block_max(static_cast<double>(mG), rmx)));
    __syncthreads();

    const TileRec tr = A.tiles[t];
    const int cnt = tr.count;
    // with UNITS the pixel moments are the unit's (k_adjoint_moments), taken
    // about the unit's origin; this tile's atoms shift to it by dsh
    int uhead = t, usize = 1;
    double uox = tr.ox, uoy = tr.oy, uoz = tr.oz;
    if (UNITS) {
        unit_of(t, A.n_tiles, A.grp_r, A.grp_rmax, uhead, usize);
        if (usize > 1) {
            const TileRec b0 = A.tiles[uhead];
            const TileRec b1 = A.tiles[uhead + 1];
            const TileRec b2 = A.tiles[usize == 4 ? uhead + 2 : uhead];
            const TileRec b3 = A.tiles[usize == 4 ? uhead + 3 : uhead];
            uox = unit_origin(usize, b0.ox, b1.ox, b2.ox, b3.ox);
            uoy = unit_origin(usize, b0.oy, b1.oy, b2.oy, b3.oy);
            uoz = unit_origin(usize, b0.oz, b1.oz, b2.oz, b3.oz);
        }
    }
    const float dshx = static_cast<float>((tr.ox - uox) * A.K);
    const float dshy = static_cast<float>((tr.oy - uoy) * A.K);
    const float dshz = static_cast<float>((tr.oz - uoz) * A.K);
    const float2* __restrict__ Gt = A.gmma + static_cast<int64_t>(tr.species) * A.npix;
    const int h = lane >> 5, q = lane & 31;
    const int Ny = A.Ny;
    const int npairs = A.n_frames * A.n_patches;
    float* rw = &red[UNITS ? 0 : wave][0];
    // this lane's atoms lane + 64 i and their FP32 partial Z
    float zr[4][3], zi[4][3];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int d = 0; d < 3; ++d) { zr[i][d] = 0.f; zi[i][d] = 0.f; }
    int since = 0;
    // every wave runs the same number of pair iterations (barriers inside)
    const int iters = (npairs + POLY_WAVES - 1) / POLY_WAVES;

    for (int it = 0; it < iters; ++it) {
        const int pr = it * POLY_WAVES + wave;
        const bool live = pr < npairs;
        const int prc = live ? pr : 0;
        const int f = prc / A.n_patches, patch = prc % A.n_patches;
        const FrameRec& F = A.frames[f];
        const PatchGeom P = patch_of(patch % A.pby, patch / A.pby, A.Ny, A.Nz);
        const double R00 = F.R[0], R01 = F.R[1], R02 = F.R[2];
        const double R10 = F.R[3], R11 = F.R[4], R12 = F.R[5];
        const double R20 = F.R[6], R21 = F.R[7], R22 = F.R[8];
        const int pmid = (P.z0 + P.mz) * Ny + P.y0 + P.my;

        float vc[3], Gy[3], Hy[3], Gz[3], Hz[3];
        double vc0[3];
        {
            const int samp[5] = {pmid, pmid - P.my, pmid - P.my + P.vy - 1, pmid - P.mz * Ny,
                                 pmid + (P.vz - 1 - P.mz) * Ny};
            double vs[5][3];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const double4 u = A.u0h[samp[k]];
                const double wx = 1.0 - u.x, wy = -u.y, wz = -u.z;
                vs[k][0] = R00 * wx + R10 * wy + R20 * wz;
                vs[k][1] = R01 * wx + R11 * wy + R21 * wz;
                vs[k][2] = R02 * wx + R12 * wy + R22 * wz;
            }
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                double gq, hq;
                quad_fit(vs[1][d] - vs[0][d], vs[2][d] - vs[0][d], -P.my, P.vy - 1 - P.my, gq, hq);
                Gy[d] = static_cast<float>(16.0 * gq);
                Hy[d] = static_cast<float>(256.0 * hq);
                quad_fit(vs[3][d] - vs[0][d], vs[4][d] - vs[0][d], -P.mz, P.vz - 1 - P.mz, gq, hq);
                Gz[d] = static_cast<float>(16.0 * gq);
                Hz[d] = static_cast<float>(256.0 * hq);
                vc0[d] = vs[0][d];
            }
        }
        const double4 um = A.u0h[pmid];
        const double Ox = fma(R00, uox, fma(R01, uoy, fma(R02, uoz, F.T[0])));
        const double Oy = fma(R10, uox, fma(R11, uoy, fma(R12, uoz, F.T[1])));
        const double Oz = fma(R20, uox, fma(R21, uoy, fma(R22, uoz, F.T[2])));
        const double O2 = Ox * Ox + Oy * Oy + Oz * Oz;
        const double uo = um.x * Ox + um.y * Oy + um.z * Oz;
        double ph_t = A.theta0[pmid] + A.K * (Ox - uo + (O2 - uo * uo) * um.w);
        ph_t -= floor(ph_t);
        const double h2 = 2.0 * um.w;
        const double cxl = (Ox - uo * um.x) * h2, cyl = (Oy - uo * um.y) * h2,
                     czl = (Oz - uo * um.z) * h2;
        vc[0] = static_cast<float>(vc0[0] + (R00 * cxl + R10 * cyl + R20 * czl));
        vc[1] = static_cast<float>(vc0[1] + (R01 * cxl + R11 * cyl + R21 * czl));
        vc[2] = static_cast<float>(vc0[2] + (R02 * cxl + R12 * cyl + R22 * czl));
        const float xc0 = static_cast<float>(F.xs[0] - vc0[0]);
        const float xc1 = static_cast<float>(F.xs[1] - vc0[1]);
        const float xc2 = static_cast<float>(F.xs[2] - vc0[2]);
        const float hh = SECOND ? static_cast<float>(um.w / A.K) : 0.f;
        const float kox = static_cast<float>(A.K * Ox), koy = static_cast<float>(A.K * Oy),
                    koz = static_cast<float>(A.K * Oz);
        const float pht = static_cast<float>(ph_t);

        float Mr[NM], Mi[NM];
        if constexpr (UNITS) {
            const float* mo = A.mom + (static_cast<int64_t>(uhead) * npairs + prc) * (2 * NM);
#pragma unroll
            for (int kk = 0; kk < NM; ++kk) {
                Mr[kk] = live ? mo[2 * kk] : 0.f;
                Mi[kk] = live ? mo[2 * kk + 1] : 0.f;
            }
        } else {
        // W on this lane's row a = q, columns b = 16 s + 8 h + u, folded into
        // H_n = sum_b W y_b^n
        float Hr[DM + 1], Hi[DM + 1];
#pragma unroll
        for (int n = 0; n <= DM; ++n) { Hr[n] = 0.f; Hi[n] = 0.f; }
        const float2* __restrict__ gf = A.gin + static_cast<int64_t>(f) * A.npix;
#pragma unroll
        for (int st = 0; st < 2; ++st) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if ((u & 3) == 0) __builtin_amdgcn_sched_barrier(0);
                const int b = 16 * st + 8 * h + u;
                const bool okp = live && q < P.vy && b < P.vz;
                const int p = (P.z0 + min(b, P.vz - 1)) * Ny + P.y0 + min(q, P.vy - 1);
                float wr = 0.f, wi = 0.f;
                if (okp) {
                    const float4 du = A.dup[p];
                    const float phs = __builtin_amdgcn_fractf(
                        pht - fmaf(koz, du.z, fmaf(koy, du.y, kox * du.x)));
                    const float sn = __builtin_amdgcn_sinf(phs), cs = __builtin_amdgcn_cosf(phs);
                    const float2 g0 = Gt[p], q0 = gf[p];
                    const float gr = g0.x * sG, gi = g0.y * sG, qr = q0.x * sg, qi = q0.y * sg;
                    const float zr_ = gr * cs - gi * sn, zi_ = gr * sn + gi * cs;
                    wr = qr * zr_ + qi * zi_;          // conj(g) z
                    wi = qr * zi_ - qi * zr_;
                }
                const float y = static_cast<float>(b - P.mz) * (1.f / 16.f);
                float yn = 1.f;
#pragma unroll
                for (int n = 0; n <= DM; ++n) {
                    Hr[n] = fmaf(wr, yn, Hr[n]);
                    Hi[n] = fmaf(wi, yn, Hi[n]);
                    yn *= y;
                }
            }
        }
        // M_mn = sum over lanes of x_q^m H_n, in two rounds through LDS
        const float xq = static_cast<float>(q - P.my) * (1.f / 16.f);
        float xm[DM + 1];
        xm[0] = 1.f;
#pragma unroll
        for (int m = 1; m <= DM; ++m) xm[m] = xm[m - 1] * xq;
#pragma unroll
        for (int round = 0; round < 2; ++round) {
            int k = 0;
#pragma unroll
            for (int m = 0; m <= DM; ++m) {
#pragma unroll
                for (int n = 0; n <= DM - m; ++n, ++k) {
                    if (k / HALF == round) {
                        const int c = k - round * HALF;
                        rw[lane * NVP + 2 * c] = xm[m] * Hr[n];
                        rw[lane * NVP + 2 * c + 1] = xm[m] * Hi[n];
                    }
                }
            }
            __syncthreads();
            const int col = min(lane, NVP - 1);
            float tot = 0.f;
#pragma unroll 8
            for (int l = 0; l < 64; ++l) tot += rw[l * NVP + col];
            __syncthreads();
#pragma unroll
            for (int kk = 0; kk < NM; ++kk) {
                if (kk / HALF == round) {
                    const int c = kk - round * HALF;
                    Mr[kk] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tot), 2 * c));
                    Mi[kk] = __int_as_float(
                        __builtin_amdgcn_readlane(__float_as_int(tot), 2 * c + 1));
                }
            }
        }
        }
        // index of moment (m, n), m + n <= DM, in the order above
        auto mk = [](int m, int n) { return m * (DM + 1) - m * (m - 1) / 2 + n; };

        // per atom: expansion coefficients, five contractions, Z
        const float Pk0 = TWO_PI_F * Gy[0], Pk1 = TWO_PI_F * Gy[1], Pk2 = TWO_PI_F * Gy[2];
        const float Rk0 = TWO_PI_F * Gz[0], Rk1 = TWO_PI_F * Gz[1], Rk2 = TWO_PI_F * Gz[2];
        const float Qk0 = TWO_PI_F * Hy[0], Qk1 = TWO_PI_F * Hy[1], Qk2 = TWO_PI_F * Hy[2];
        const float Sk0 = TWO_PI_F * Hz[0], Sk1 = TWO_PI_F * Hz[1], Sk2 = TWO_PI_F * Hz[2];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const AtomRec a = tile[lane + 64 * i];
            const float sx = a.sx + dshx, sy = a.sy + dshy, sz = a.sz + dshz;
            float e0 = fmaf(vc[2], sz, fmaf(vc[1], sy, vc[0] * sx));
            if (SECOND) {
                const float w = fmaf(xc2, sz, fmaf(xc1, sy, xc0 * sx)) - e0;
                const float s2 = UNITS ? fmaf(sz, sz, fmaf(sy, sy, sx * sx)) : a.s2;
                e0 = fmaf(hh, fmaf(-w, w, s2), e0);
            }
            const float Pj = fmaf(Pk2, sz, fmaf(Pk1, sy, Pk0 * sx));
            const float Rj = fmaf(Rk2, sz, fmaf(Rk1, sy, Rk0 * sx));
            const float Qj = fmaf(Qk2, sz, fmaf(Qk1, sy, Qk0 * sx));
            const float Sj = fmaf(Sk2, sz, fmaf(Sk1, sy, Sk0 * sx));
            float am[D + 1], bn[D + 1];
            am[0] = 1.f;
            bn[0] = 1.f;
#pragma unroll
            for (int m = 1; m <= D; ++m) {
                am[m] = am[m - 1] * Pj * (1.f / m);
                bn[m] = bn[m - 1] * Rj * (1.f / m);
            }
            // S_k = sum_mn c_mn M_{(m,n) + shift_k}, shifts (0,0) (1,0) (2,0) (0,1) (0,2),
            // with c_mn = i^{m+n} (T - i U) = i^{m+n} T + i^{m+n+3} U and T, U
            // real: each power of i swaps or negates the moment's parts, so a
            // term costs two FMAs for T and two more only where U is not zero
            float sr[5] = {0.f, 0.f, 0.f, 0.f, 0.f}, si[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
            // (ar, ai) += w i^e (Mr[kk], Mi[kk])
            auto mac = [&](float w, int e, int kk, float& ar, float& ai) {
                switch (e & 3) {
                    case 0: ar = fmaf(w, Mr[kk], ar); ai = fmaf(w, Mi[kk], ai); break;
                    case 1: ar = fmaf(-w, Mi[kk], ar); ai = fmaf(w, Mr[kk], ai); break;
                    case 2: ar = fmaf(-w, Mr[kk], ar); ai = fmaf(-w, Mi[kk], ai); break;
                    default: ar = fmaf(w, Mi[kk], ar); ai = fmaf(-w, Mr[kk], ai); break;
                }
            };
#pragma unroll
            for (int m = 0; m <= D; ++m) {
#pragma unroll
                for (int n = 0; n <= D - m; ++n) {
                    const float T = am[m] * bn[n];
                    const bool qterm = m >= 2 && m - 2 + n <= 3;
                    const bool sterm = n >= 2 && m + n - 2 <= 3;
                    float U = 0.f;
                    if (qterm) U = fmaf(Qj, am[m - 2] * bn[n], U);
                    if (sterm) U = fmaf(Sj, am[m] * bn[n - 2], U);
                    const int k0 = mk(m, n), k1 = mk(m + 1, n), k2 = mk(m + 2, n),
                              k3 = mk(m, n + 1), k4 = mk(m, n + 2);
                    const int ks[5] = {k0, k1, k2, k3, k4};
                    // the gradient-direction terms scale as |Gy|, |Gz| ~ 1e-2
                    // and |Hy|, |Hz| ~ 1e-5 of |vc|: their expansions stop at
                    // degree 4 and 2 (below 1e-7 of Z)
                    const int deg[5] = {D, 4, 2, 4, 2};
#pragma unroll
                    for (int c = 0; c < 5; ++c) {
                        if (m + n > deg[c]) continue;
                        mac(T, m + n, ks[c], sr[c], si[c]);
                        if (qterm || sterm) mac(U, m + n + 3, ks[c], sr[c], si[c]);
                    }
                }
            }
            const float ce = __builtin_amdgcn_cosf(e0), se = __builtin_amdgcn_sinf(e0);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const float tr_ = vc[d] * sr[0] + Gy[d] * sr[1] + Hy[d] * sr[2] + Gz[d] * sr[3] +
                                  Hz[d] * sr[4];
                const float ti_ = vc[d] * si[0] + Gy[d] * si[1] + Hy[d] * si[2] + Gz[d] * si[3] +
                                  Hz[d] * si[4];
                zr[i][d] += ce * tr_ - se * ti_;
                zi[i][d] += ce * ti_ + se * tr_;
            }
        }
        if (++since == 16 || it + 1 == iters) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int j = lane + 64 * i;
                if (j < cnt) {
#pragma unroll
                    for (int d = 0; d < 3; ++d) {
                        atomicAdd(&zacc[j][2 * d], static_cast<double>(zr[i][d]));
                        atomicAdd(&zacc[j][2 * d + 1], static_cast<double>(zi[i][d]));
                    }
                }
#pragma unroll
                for (int d = 0; d < 3; ++d) { zr[i][d] = 0.f; zi[i][d] = 0.f; }
            }
            since = 0;
        }
    }
    __syncthreads();
    if (tid < cnt) {
        const AtomRec a = tile[tid];
        const double scale = -2.0 * PI * A.K * 1e-10 /
                             (static_cast<double>(sg) * static_cast<double>(sG));
        const uint32_t idx = __float_as_uint(a.pad0);
        double wi, wr;
        sincospi(2.0 * static_cast<double>(a.ph), &wi, &wr);
        wr *= a.mag;
        wi *= a.mag;
        for (int c = 0; c < 3; ++c) {
            const double g = scale * (wr * zacc[tid][2 * c + 1] + wi * zacc[tid][2 * c]);
            A.grad[3 * static_cast<int64_t>(idx) + c] += static_cast<float>(g);
        }
    }
}

// Sum slice partials in slice order and add into the output field.
template <bool C128>
__global__ void k_reduce(const double2* __restrict__ partial, int n_slices, int npix,
                         int n_frames, void* __restrict__ out) {
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= static_cast<int64_t>(npix) * n_frames) return;
    const int64_t f = q / npix, p = q % npix;
    const double2* src = partial + f * n_slices * (int64_t)npix + p;
    double r = 0.0, im = 0.0;
    for (int s = 0; s < n_slices; ++s) {
        const double2 v = src[static_cast<int64_t>(s) * npix];
        r += v.x;
        im += v.y;
    }
    if (C128) {
        double2* o = static_cast<double2*>(out) + q;
        double2 cur = *o;
        cur.x += r;
        cur.y += im;
        *o = cur;
    } else {
        float2* o = static_cast<float2*>(out) + q;
        float2 cur = *o;
        cur.x += static_cast<float>(r);
        cur.y += static_cast<float>(im);
        *o = cur;
    }
}

// ------------------------------------------------------------------ Ein
__global__ void k_ein(const float* __restrict__ pos, int64_t n, float R00, float R01, float R02,
                      float R10, float R11, float R12, float R20, float R21, float R22,
                      float T0, float T1, float T2, const float2* __restrict__ E0, int NyB,
                      int NzB, double du, double dv, double uc, double vc,
                      float2* __restrict__ out) {
    // Same float32 roundings, in the same order, as the reference's NumPy
    // sampler: plain operators with multiply-add contraction off.
#pragma clang fp contract(off)
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
    // lab y and z; e1 = +y, e2 = +z for a +x beam
    const float ly = ((x * R10 + y * R11) + z * R12) + T1;
    const float lz = ((x * R20 + y * R21) + z * R22) + T2;
    (void)R00; (void)R01; (void)R02; (void)T0;
    const float iu = ly / static_cast<float>(du) + static_cast<float>(uc);
    const float iv = lz / static_cast<float>(dv) + static_cast<float>(vc);
    float2 res = make_float2(0.f, 0.f);
    if (iu >= 0.f && iu <= static_cast<float>(NyB - 1) && iv >= 0.f &&
        iv <= static_cast<float>(NzB - 1)) {
        const int i0 = static_cast<int>(floorf(iu));
        const int j0 = static_cast<int>(floorf(iv));
        const int i1 = min(i0 + 1, NyB - 1);
        const int j1 = min(j0 + 1, NzB - 1);
        const float fu = iu - static_cast<float>(i0);
        const float fv = iv - static_cast<float>(j0);
        const float w00 = (1.f - fu) * (1.f - fv);
        const float w01 = (1.f - fu) * fv;
        const float w10 = fu * (1.f - fv);
        const float w11 = fu * fv;
        const float2 a = E0[i0 * NzB + j0], b = E0[i0 * NzB + j1];
        const float2 c = E0[i1 * NzB + j0], d = E0[i1 * NzB + j1];
        res.x = ((a.x * w00 + b.x * w01) + c.x * w10) + d.x * w11;
        res.y = ((a.y * w00 + b.y * w01) + c.y * w10) + d.y * w11;
    }
    out[i] = res;
}

template <int PIX, bool SECOND>
hipError_t launch_sum(const SumArgs& A, int pix_blocks, int n_slices, int n_frames,
                      hipStream_t st) {
    dim3 grid(pix_blocks, n_slices, n_frames);
    hipLaunchKernelGGL((k_sum<PIX, SECOND>), grid, dim3(SUM_THREADS), 0, st, A);
    return hipGetLastError();
}

template <bool POL, int ANA>
hipError_t launch_exact(const ExactArgs& A, int n_slices, int n_frames, hipStream_t st) {
    constexpr int PIX = 4;
    dim3 grid(blocks_for(A.npix, SUM_THREADS * PIX), n_slices, n_frames);
    hipLaunchKernelGGL((k_sum_exact<PIX, POL, ANA>), grid, dim3(SUM_THREADS), 0, st, A);
    return hipGetLastError();
}

}  // namespace
}  // namespace mosaic

using namespace mosaic;

extern "C" int mosaic_kin_workspace(int64_t atom_count, int n_species, const int64_t* counts,
                                    int n_frames, int Ny, int Nz, size_t* bytes) {
    if (!bytes || atom_count < 0 || n_species < 1 || n_species > MAX_SPECIES || !counts ||
        n_frames < 1 || n_frames > MAX_FRAMES || Ny < 1 || Nz < 1 ||
        static_cast<int64_t>(Ny) * Nz > MAX_PIXELS ||
        static_cast<int64_t>(Ny) * Nz * n_frames > MAX_OUTPUTS ||
        atom_count >= (int64_t(1) << 32))
        return MOSAIC_ERR_ARGS;
    int64_t tot = 0;
    for (int s = 0; s < n_species; ++s) {
        if (counts[s] < 0) return MOSAIC_ERR_ARGS;
        tot += counts[s];
    }
    if (tot != atom_count) return MOSAIC_ERR_ARGS;
    Layout L;
    const int rc = make_layout(nullptr, atom_count, n_species, counts, n_frames, Ny, Nz, L);
    if (rc != MOSAIC_OK) return rc;
    *bytes = L.total;
    return MOSAIC_OK;
}

// Page-locked host staging for the staging call's tables (the k-d plan's
// nodes and chunks, the in-LDS nodes) and its read-back, one per host thread:
// copies from pageable memory stall the stream.  An event after the last
// copy that reads the buffer guards its reuse.
namespace {
struct PinnedStage {
    char* buf = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
    bool pending = false;
};
thread_local PinnedStage g_pinned;

// at least `bytes` of page-locked memory whose previous copies have finished
char* pinned_stage(size_t bytes) {
    PinnedStage& P = g_pinned;
    if (P.pending) {
        if (hipEventSynchronize(P.done) != hipSuccess) return nullptr;
        P.pending = false;
    }
    if (!P.done && hipEventCreateWithFlags(&P.done, hipEventDisableTiming) != hipSuccess) {
        P.done = nullptr;
        return nullptr;
    }
    if (bytes > P.cap) {
        if (P.buf) (void)hipHostFree(P.buf);
        P.buf = nullptr;
        P.cap = 0;
        const size_t want = std::max<size_t>(bytes, size_t(1) << 20);
        if (hipHostMalloc(reinterpret_cast<void**>(&P.buf), want, hipHostMallocDefault) != hipSuccess)
            return nullptr;
        P.cap = want;
    }
    return P.buf;
}

void pinned_stage_mark(hipStream_t st) {
    if (g_pinned.done && hipEventRecord(g_pinned.done, st) == hipSuccess) g_pinned.pending = true;
}
}  // namespace

extern "C" int mosaic_kin_stage(int64_t atom_count, const float* pos, const uint8_t* species,
                                const float* amp, int n_species, const int64_t* counts, double k,
                                int n_frames, int Ny, int Nz, void* workspace,
                                size_t workspace_bytes, double* stage_stats,
                                mosaic_stream_t stream) {
    size_t need = 0;
    int rc = mosaic_kin_workspace(atom_count, n_species, counts, n_frames, Ny, Nz, &need);
    if (rc != MOSAIC_OK) return rc;
    if (!workspace || workspace_bytes < need) return MOSAIC_ERR_WORKSPACE;
    if (atom_count > 0 && (!pos || !species)) return MOSAIC_ERR_ARGS;
    if (!(k > 0.0)) return MOSAIC_ERR_ARGS;
    hipStream_t st = static_cast<hipStream_t>(stream);
    Layout L;
    rc = make_layout(workspace, atom_count, n_species, counts, n_frames, Ny, Nz, L);
    if (rc != MOSAIC_OK) return rc;
    if (stage_stats) {
        stage_stats[0] = 0.0;
        stage_stats[1] = 0.0;
    }
    if (atom_count == 0) return MOSAIC_OK;

    SpeciesSplit split;
    split.n_species = n_species;
    split.atom_start[0] = 0;
    split.tile_start[0] = 0;
    for (int s = 0; s < n_species; ++s) {
        split.atom_start[s + 1] = split.atom_start[s] + counts[s];
        split.tile_start[s + 1] =
            split.tile_start[s] + static_cast<int>((counts[s] + TILE - 1) / TILE);
    }
    const int glevels = kd_global_levels(n_species, counts);
    // the host plan first (host tables alive until the synchronisation
    // below), so that it overlaps the stream's earlier work
    std::vector<KdgNode> gnodes;
    std::vector<KdgChunk> gchunks;
    std::vector<size_t> node_off, chunk_off;
    kdg_tables(n_species, counts, gnodes, gchunks, node_off, chunk_off);
    std::vector<std::vector<int64_t>> RB, RE;
    kd_plan_levels(n_species, counts, glevels, RB, RE);
    const std::vector<int64_t>& nbeg = RB[glevels];
    const std::vector<int64_t>& nend = RE[glevels];
    // nodes of the in-LDS finish, grouped by size: up to 8 tiles, 16, 32
    std::vector<KdNode> lnodes;
    int lcount[3] = {0, 0, 0};
    for (int cls = 0; cls < 3; ++cls) {
        for (size_t h = 0; h < nbeg.size(); ++h) {
            const int64_t cnt = nend[h] - nbeg[h];
            if (cnt <= 0) continue;
            if (cnt > KDL_MAX) return MOSAIC_ERR_ARGS;   // not reached by the plan
            if ((cnt <= (TILE << 3) ? 0 : cnt <= (TILE << 4) ? 1 : 2) != cls) continue;
            const int s = static_cast<int>(h >> glevels);
            KdNode nd;
            nd.begin = static_cast<uint32_t>(nbeg[h]);
            nd.count = static_cast<uint32_t>(cnt);
            nd.tile0 = static_cast<uint32_t>(split.tile_start[s] +
                                             (nbeg[h] - split.atom_start[s]) / TILE);
            nd.species = static_cast<uint32_t>(s);
            lnodes.push_back(nd);
            ++lcount[cls];
        }
    }
    if (lnodes.size() > static_cast<size_t>(std::max(1, L.n_tiles))) return MOSAIC_ERR_ARGS;
    hipLaunchKernelGGL(k_stage_init, dim3(1), dim3(64), 0, st, L.radius_bits);
    // group by species (stable), then bisect each species level by level:
    // the global levels by chunks, the last (at most KDL_LEVELS) in LDS.
    // The grouping is a counting sort straight into the first coordinate set
    // (its index array holds the order); the levels then move the
    // coordinates between two sets
    KdgSet cur = {L.kdg_xyz[0][0], L.kdg_xyz[0][1], L.kdg_xyz[0][2], L.vals_b};
    KdgSet other = {L.kdg_xyz[1][0], L.kdg_xyz[1][1], L.kdg_xyz[1][2], L.vals_a};
    {
        const int nblk = static_cast<int>((atom_count + KDG_CHUNK - 1) / KDG_CHUNK);
        hipLaunchKernelGGL(k_species_count, dim3(nblk), dim3(KDG_THREADS), 0, st, species, atom_count,
                           L.spc_cnt);
        hipLaunchKernelGGL(k_species_prefix, dim3(n_species), dim3(KDG_THREADS), 0, st, L.spc_cnt, nblk,
                           split, L.spc_off);
        hipLaunchKernelGGL(k_species_scatter, dim3(nblk), dim3(KDG_THREADS), 0, st, pos, species,
                           atom_count, n_species, L.spc_off, cur);
        MOSAIC_HIP_CHECK(hipGetLastError());
    }
    // the plan's tables (global nodes, chunks, in-LDS nodes) through
    // page-locked memory laid out as in the workspace, one upload
    const size_t gn_b = gnodes.size() * sizeof(KdgNode), gc_b = gchunks.size() * sizeof(KdgChunk);
    const size_t ln_b = lnodes.size() * sizeof(KdNode);
    const size_t off_c = static_cast<size_t>(reinterpret_cast<char*>(L.kdg_chunks) -
                                              reinterpret_cast<char*>(L.kdg_nodes));
    const size_t off_l = static_cast<size_t>(reinterpret_cast<char*>(L.kd_nodes) -
                                              reinterpret_cast<char*>(L.kdg_nodes));
    const size_t span = off_l + ln_b;
    char* pin = pinned_stage((span + 255) / 256 * 256 + 256);
    if (!pin) return MOSAIC_ERR_HIP;
    std::memcpy(pin, gnodes.data(), gn_b);
    std::memcpy(pin + off_c, gchunks.data(), gc_b);
    std::memcpy(pin + off_l, lnodes.data(), ln_b);
    unsigned int* pin_bits = reinterpret_cast<unsigned int*>(pin + (span + 255) / 256 * 256);
    MOSAIC_HIP_CHECK(hipMemcpyAsync(L.kdg_nodes, pin, span, hipMemcpyHostToDevice, st));
    pinned_stage_mark(st);
    if (glevels > 0) {
        auto tight = [&](int lv) { return lv == 0 || lv >= glevels - KD_TIGHT_LEVELS; };
        for (int l = 0; l < glevels; ++l) {
            const int64_t nodes = static_cast<int64_t>(node_off[l + 1] - node_off[l]);
            const unsigned int nchunks = static_cast<unsigned int>(chunk_off[l + 1] - chunk_off[l]);
            const KdgNode* nd = L.kdg_nodes + node_off[l];
            const KdgChunk* ch = L.kdg_chunks + chunk_off[l];
            float* fbox = L.kd_fbox[l & 1];
            float* fnext = (l + 1 < glevels && !tight(l + 1)) ? L.kd_fbox[(l + 1) & 1] : nullptr;
            if (nchunks == 0) continue;
            if (tight(l)) {
                hipLaunchKernelGGL(k_kdg_box, dim3(nchunks), dim3(KDG_THREADS), 0, st, cur, ch,
                                   L.kdg_cbox);
                hipLaunchKernelGGL(k_kdg_box_node, dim3(static_cast<unsigned int>(nodes)), dim3(64), 0,
                                   st, nd, L.kdg_cbox, fbox);
            }
            hipLaunchKernelGGL(k_kdg_bin, dim3(nchunks), dim3(KDG_THREADS), 0, st, cur, ch, nd,
                               fbox, L.kdg_q8, L.kdg_bh);
            const int S = kdg_slices(gnodes, node_off[l], node_off[l + 1]);
            hipLaunchKernelGGL(k_kdg_total, dim3(S, static_cast<unsigned int>(nodes)),
                               dim3(KDG_THREADS), 0, st, nd, L.kdg_bh, S, L.kdg_part);
            hipLaunchKernelGGL(k_kdg_select, dim3(static_cast<unsigned int>(nodes)),
                               dim3(KDG_THREADS), 0, st, nd, L.kdg_part, S, L.kdg_bh, fbox, fnext,
                               L.kdg_sel, L.kdg_boff);
            hipLaunchKernelGGL(k_kdg_scatter, dim3(nchunks), dim3(KDG_THREADS), 0, st, ch, nd,
                               L.kdg_q8, cur, L.kdg_sel, L.kdg_boff, other);
            MOSAIC_HIP_CHECK(hipGetLastError());
            std::swap(cur, other);
        }
    }
    const double K = k * INV_2PI_HI;
    if (!lnodes.empty()) {
        const float2* a2 = reinterpret_cast<const float2*>(amp);
        const KdNode* nds = L.kd_nodes;
        // the final order goes to vals_b (mosaic_kin_restage reads it there)
        if (lcount[0] > 0)
            hipLaunchKernelGGL(k_kd_local<3>, dim3(lcount[0]), dim3(32 << 3), 0, st, cur, a2,
                               L.vals_b, nds, K, L.tiles, L.atoms, L.grp_r, L.tile_pos,
                               L.radius_bits);
        if (lcount[1] > 0)
            hipLaunchKernelGGL(k_kd_local<4>, dim3(lcount[1]), dim3(32 << 4), 0, st, cur, a2,
                               L.vals_b, nds + lcount[0], K, L.tiles, L.atoms, L.grp_r,
                               L.tile_pos, L.radius_bits);
        if (lcount[2] > 0)
            hipLaunchKernelGGL(k_kd_local<5>, dim3(lcount[2]), dim3(32 << 5), 0, st, cur, a2,
                               L.vals_b, nds + lcount[0] + lcount[1], K, L.tiles, L.atoms,
                               L.grp_r, L.tile_pos, L.radius_bits);
    }
    MOSAIC_HIP_CHECK(hipGetLastError());
    if (!stage_stats) return MOSAIC_OK;             // left in the workspace, no wait
    MOSAIC_HIP_CHECK(hipMemcpyAsync(pin_bits, L.radius_bits, 2 * sizeof(unsigned int),
                                    hipMemcpyDeviceToHost, st));
    pinned_stage_mark(st);
    MOSAIC_HIP_CHECK(hipStreamSynchronize(st));
    float rf[2];
    std::memcpy(rf, pin_bits, sizeof(rf));
    stage_stats[0] = static_cast<double>(rf[0]);
    stage_stats[1] = static_cast<double>(rf[1]);
    return MOSAIC_OK;
}

extern "C" int mosaic_kin_restage(int64_t atom_count, const float* pos, const float* amp,
                                  int n_species, const int64_t* counts, double k, int n_frames,
                                  int Ny, int Nz, void* workspace, size_t workspace_bytes,
                                  double* stage_stats, mosaic_stream_t stream) {
    size_t need = 0;
    int rc = mosaic_kin_workspace(atom_count, n_species, counts, n_frames, Ny, Nz, &need);
    if (rc != MOSAIC_OK) return rc;
    if (!workspace || workspace_bytes < need) return MOSAIC_ERR_WORKSPACE;
    if (atom_count > 0 && !pos) return MOSAIC_ERR_ARGS;
    if (!(k > 0.0) || !stage_stats) return MOSAIC_ERR_ARGS;
    hipStream_t st = static_cast<hipStream_t>(stream);
    Layout L;
    rc = make_layout(workspace, atom_count, n_species, counts, n_frames, Ny, Nz, L);
    if (rc != MOSAIC_OK) return rc;
    stage_stats[0] = 0.0;
    stage_stats[1] = 0.0;
    if (atom_count == 0) return MOSAIC_OK;
    SpeciesSplit split;
    split.n_species = n_species;
    split.atom_start[0] = 0;
    split.tile_start[0] = 0;
    for (int s = 0; s < n_species; ++s) {
        split.atom_start[s + 1] = split.atom_start[s] + counts[s];
        split.tile_start[s + 1] =
            split.tile_start[s] + static_cast<int>((counts[s] + TILE - 1) / TILE);
    }
    hipLaunchKernelGGL(k_stage_init, dim3(1), dim3(64), 0, st, L.radius_bits);
    // the order of the last mosaic_kin_stage on this workspace is in vals_b
    hipLaunchKernelGGL(k_tiles, dim3(L.n_tiles), dim3(TILE), 0, st, pos,
                       reinterpret_cast<const float2*>(amp), L.vals_b, split, k * INV_2PI_HI,
                       L.tiles, L.atoms, L.radius_bits);
    // the units' radii from the new records (the node places come from the
    // k-d staging of this workspace)
    hipLaunchKernelGGL(k_unit_init, dim3(blocks_for(L.n_tiles, 256)), dim3(256), 0, st,
                       L.tile_pos, L.n_tiles, L.grp_r);
    hipLaunchKernelGGL(k_unit_radii, dim3(L.n_tiles), dim3(TILE), 0, st, L.tiles, L.atoms,
                       L.tile_pos, k * INV_2PI_HI, L.grp_r);
    MOSAIC_HIP_CHECK(hipGetLastError());
    unsigned int bits[2] = {0u, 0u};
    MOSAIC_HIP_CHECK(hipMemcpyAsync(bits, L.radius_bits, sizeof(bits), hipMemcpyDeviceToHost, st));
    MOSAIC_HIP_CHECK(hipStreamSynchronize(st));
    float rf[2];
    std::memcpy(rf, bits, sizeof(rf));
    stage_stats[0] = static_cast<double>(rf[0]);
    stage_stats[1] = static_cast<double>(rf[1]);
    return MOSAIC_OK;
}

namespace {
void fill_species(const mosaic_species_t* tab, int n, SpeciesConst* out) {
    for (int s = 0; s < n; ++s) {
        for (int m = 0; m < 5; ++m) {
            out[s].a[m] = tab[s].wk[m];
            out[s].b[m] = tab[s].wk[6 + m];
        }
        out[s].c = tab[s].wk[5];
        out[s].f0z = tab[s].f0_zero;
        out[s].fp = tab[s].fp;
        out[s].fpp = tab[s].fpp;
    }
}
}  // namespace

extern "C" int mosaic_kin_pixels(int n_species, const mosaic_species_t* species_tab,
                                 const mosaic_kin_opts_t* opts, int Ny, int Nz,
                                 const double* coords, void* workspace, size_t workspace_bytes,
                                 double* pixel_geom, mosaic_stream_t stream) {
    if (Ny < 1 || Nz < 1 || static_cast<int64_t>(Ny) * Nz > MAX_PIXELS || !coords || !opts ||
        !(opts->k > 0.0) || !workspace || n_species < 1 ||
        n_species > MAX_SPECIES || !species_tab)
        return MOSAIC_ERR_ARGS;
    const int npix = Ny * Nz;
    Carver c(workspace);
    PixelTables T;
    take_pixel_tables(c, npix, n_species, T);
    if (c.off > workspace_bytes) return MOSAIC_ERR_WORKSPACE;
    PixOpts O;
    O.k = opts->k;
    O.pol_deg = opts->pol_deg;
    for (int d = 0; d < 3; ++d) O.pol_e[d] = opts->pol_e[d];
    O.remove_forward = opts->remove_forward ? 1 : 0;
    O.apply_pol = opts->apply_pol ? 1 : 0;
    O.n_species = n_species;
    fill_species(species_tab, n_species, O.spc);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int pb = blocks_for(npix, PRO_THREADS);
    hipLaunchKernelGGL(k_pixels, dim3(pb), dim3(PRO_THREADS), 0, st, coords, coords + npix,
                       coords + 2 * (int64_t)npix, Ny, Nz, O, T);
    hipLaunchKernelGGL(k_stats_reduce, dim3(1), dim3(256), 0, st, T.stat_blocks, pb, T.stats);
    MOSAIC_HIP_CHECK(hipGetLastError());
    if (!pixel_geom) return MOSAIC_OK;              // left in the workspace, no wait
    MOSAIC_HIP_CHECK(hipMemcpyAsync(pixel_geom, T.stats, N_GEOM * sizeof(double),
                                    hipMemcpyDeviceToHost, st));
    MOSAIC_HIP_CHECK(hipStreamSynchronize(st));
    return MOSAIC_OK;
}

namespace {
// Row-recurrence bound: the deviation of v from the fitted line over KP
// pixels (second difference d2u), times the largest offset, in revolutions.
double rec_curv_bound(int kp, double d2u, double radius_wl) {
    const double h = 0.5 * (kp - 1);
    return 0.5 * (h * h / 2.0) * d2u * radius_wl;
}
}  // namespace

// Degree of the moment expansion (k_sum_poly, k_adjoint_poly), 0 where it
// does not apply: patch_ok carries the direction model's bounds (the
// separable model's for the adjoint, the model with the cross term for the
// sum, which also passes cross to bound the cross term's truncation), plus
// the truncation.
// phi bounds |P x + R y| over a patch from the largest first differences
// along a row and a column (16 pixels each way, with the curvature); qmax
// bounds the quadratic model terms |Q|, |S|.  Both truncations are per-atom
// bounds at the patch corner for the largest tile, each within 1e-6 of the
// atom's term (the FP32 offset phase already costs ~3e-6 of the field).
// The moment sum also takes the per-(tile, pixel) origin phase,
// K |O| |u_p - u_mid| < K omax 16 sqrt(2) (du_row + du_col) revolutions,
// into v_sin / v_cos without reduction, so that must stay within their range.
static int poly_degree(bool patch_ok, double radius_wl, const double* pixel_geom, double k_omax,
                       bool cross = false) {
    if (!patch_ok) return 0;
    const double du1 = pixel_geom[1], du2 = pixel_geom[2];
    const double du2z = pixel_geom[4], du1z = pixel_geom[5];
    if (k_omax * 16.0 * 1.4142135623730951 * (du1 + du1z + 16.0 * (du2 + du2z)) + 1.0 > 200.0)
        return 0;
    const double phi = 2.0 * PI * 16.0 * ((du1 + 16.0 * du2) + (du1z + 16.0 * du2z)) * radius_wl;
    const double qmax = 2.0 * PI * 256.0 * 0.5 * std::max(du2, du2z) * radius_wl;
    if (qmax > 3e-4 || qmax * phi * phi * phi * phi / 24.0 > 1e-6) return 0;
    if (cross) {   // the cross term, kept where the rest of its monomial has degree <= 1
        const double umax = 2.0 * PI * 256.0 * pixel_geom[7] * radius_wl;
        if (umax > 3e-4 || umax * phi * phi / 2.0 > 1e-6) return 0;
    }
    auto trunc = [&](int D) {
        double t = 1.0;
        for (int i = 1; i <= D + 1; ++i) t *= phi / i;
        return t;
    };
    if (trunc(5) <= 1e-6) return 5;
    if (trunc(6) <= 1e-6) return 6;
    return 0;
}

// Bounds of the matrix-core moment sum (k_gm_moments) for a unit of radius
// r [wavelengths]: the degree-6 expansion of k_sum_poly, with each patch's
// row and column gradients replaced by the mean over its region of 8 x 8
// patches.  The region's patch middles lie within 112 pixels (per axis) of
// their mean, so a patch's row gradient differs from the region's by at most
// 112 (du2 + |gyz|) per pixel (column: 112 (du2z + |gyz|)).  With the quadratic and cross terms
// that bounds eps, the largest |eps_j(x, y)| over a patch, which the
// expansion keeps to second order; the truncations are phi^7 / 7! (degree 6),
// eps phi^5 / 5! (first order to degree 4), eps^2 / 2 phi^2 / 2 (second
// order to degree 1) and eps^3 / 6, each within 1e-6 of the atom's term.
static bool gm_bounds(double r, const double* pixel_geom) {
    const double du1 = pixel_geom[1], du2 = pixel_geom[2];
    const double du2z = pixel_geom[4], du1z = pixel_geom[5], gyz = pixel_geom[7];
    const double phi = 2.0 * PI * 16.0 * ((du1 + 16.0 * du2) + (du1z + 16.0 * du2z)) * r;
    const double dg = 112.0 * (du2 + du2z + 2.0 * gyz);
    const double eps = 2.0 * PI * (16.0 * dg + 256.0 * (0.5 * du2 + 0.5 * du2z + gyz)) * r;
    const double p2 = phi * phi, p5 = p2 * p2 * phi, p7 = p5 * p2;
    return p7 / 5040.0 <= 1e-6 && eps * p5 / 120.0 <= 1e-6 &&
           0.25 * eps * eps * p2 <= 1e-6 && eps * eps * eps / 6.0 <= 1e-6;
}

// Launch timing of the matrix-core moment sum, for measurement (enabled by
// mosaic_kin_timing): HIP events on the sum's stream around each round's
// moments and evaluation kernels of the first TIMING_CALLS calls since timing
// was switched on, and a count of the (unit, patch) evaluations taken pixel by
// pixel over those calls, per device.  Nothing waits for the events until
// they are read.
namespace {
constexpr int TIMING_DEVICES = 16;
constexpr int TIMING_CALLS = 16;
constexpr int TIMING_ROUNDS = 16;
struct KinTiming {
    bool on = false;
    bool made = false;
    int calls = 0;                         // calls recorded since timing was switched on
    int rounds[TIMING_CALLS];              // rounds of each (0: more than TIMING_ROUNDS, not timed)
    hipEvent_t ev[TIMING_CALLS][TIMING_ROUNDS][4];
    unsigned int* fallbacks = nullptr;     // device counter since timing was switched on
};
KinTiming g_timing[TIMING_DEVICES];

KinTiming* timing_state() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= TIMING_DEVICES) return nullptr;
    return &g_timing[dev];
}

// the recording slot of a new call, or -1
int timing_slot(KinTiming* T) {
    return (T && T->on && T->made && T->calls < TIMING_CALLS) ? T->calls : -1;
}

// event k (0, 1: moments start, stop; 2, 3: evaluation) of round r of slot c
void timing_mark(KinTiming* T, int c, int r, int k, hipStream_t st) {
    if (c >= 0 && r < TIMING_ROUNDS) (void)hipEventRecord(T->ev[c][r][k], st);
}

void timing_done(KinTiming* T, int c, int rounds) {
    if (c < 0) return;
    T->rounds[c] = rounds <= TIMING_ROUNDS ? rounds : 0;
    T->calls = c + 1;
}
}  // namespace

namespace {
// The stage statistics (largest tile radius [wavelengths], largest origin
// distance [m]) and the pixel geometry that mosaic_kin_stage and
// mosaic_kin_pixels left in workspace L, into host ss[2] / pg[N_GEOM]
// (either may be null): one read through page-locked memory, one wait.
int read_stats(const Layout& L, double* ss, double* pg, hipStream_t st) {
    char* pin = pinned_stage(64 + N_GEOM * sizeof(double));
    if (!pin) return MOSAIC_ERR_HIP;
    if (ss)
        MOSAIC_HIP_CHECK(hipMemcpyAsync(pin, L.radius_bits, 2 * sizeof(unsigned int),
                                        hipMemcpyDeviceToHost, st));
    if (pg)
        MOSAIC_HIP_CHECK(hipMemcpyAsync(pin + 64, L.px.stats, N_GEOM * sizeof(double),
                                        hipMemcpyDeviceToHost, st));
    pinned_stage_mark(st);
    MOSAIC_HIP_CHECK(hipStreamSynchronize(st));
    if (ss) {
        float rf[2];
        std::memcpy(rf, pin, sizeof(rf));
        ss[0] = static_cast<double>(rf[0]);
        ss[1] = static_cast<double>(rf[1]);
    }
    if (pg) std::memcpy(pg, pin + 64, N_GEOM * sizeof(double));
    return MOSAIC_OK;
}
}  // namespace

extern "C" int mosaic_kin_stats(int64_t atom_count, int n_species, const int64_t* counts,
                                int n_frames, int Ny, int Nz, void* workspace,
                                size_t workspace_bytes, double* stage_stats, double* pixel_geom,
                                mosaic_stream_t stream) {
    size_t need = 0;
    int rc = mosaic_kin_workspace(atom_count, n_species, counts, n_frames, Ny, Nz, &need);
    if (rc != MOSAIC_OK) return rc;
    if (!workspace || workspace_bytes < need) return MOSAIC_ERR_WORKSPACE;
    Layout L;
    rc = make_layout(workspace, atom_count, n_species, counts, n_frames, Ny, Nz, L);
    if (rc != MOSAIC_OK) return rc;
    double* ss = stage_stats;
    if (ss && (atom_count == 0 || L.n_tiles == 0)) {   // nothing staged: no statistics written
        ss[0] = 0.0;
        ss[1] = 0.0;
        ss = nullptr;
    }
    if (!ss && !pixel_geom) return MOSAIC_OK;
    return read_stats(L, ss, pixel_geom, static_cast<hipStream_t>(stream));
}

extern "C" int mosaic_kin_timing(int enable, float* ms_out) {   // ms_out: host [3]
    KinTiming* T = timing_state();
    if (!T) return MOSAIC_ERR_HIP;
    if (enable && !T->made) {
        for (int c = 0; c < TIMING_CALLS; ++c)
            for (int r = 0; r < TIMING_ROUNDS; ++r)
                for (int k = 0; k < 4; ++k) MOSAIC_HIP_CHECK(hipEventCreate(&T->ev[c][r][k]));
        MOSAIC_HIP_CHECK(hipMalloc(reinterpret_cast<void**>(&T->fallbacks), sizeof(unsigned int)));
        T->made = true;
    }
    if (enable && !T->on) {                // a fresh record
        T->calls = 0;
        MOSAIC_HIP_CHECK(hipMemset(T->fallbacks, 0, sizeof(unsigned int)));
        MOSAIC_HIP_CHECK(hipDeviceSynchronize());
    }
    if (ms_out) {
        ms_out[0] = 0.f;
        ms_out[1] = 0.f;
        ms_out[2] = 0.f;
        int timed = 0;
        for (int c = 0; c < T->calls; ++c) {
            if (T->rounds[c] <= 0) continue;
            for (int r = 0; r < T->rounds[c]; ++r) {
                MOSAIC_HIP_CHECK(hipEventSynchronize(T->ev[c][r][3]));
                float a = 0.f, b = 0.f;
                MOSAIC_HIP_CHECK(hipEventElapsedTime(&a, T->ev[c][r][0], T->ev[c][r][1]));
                MOSAIC_HIP_CHECK(hipEventElapsedTime(&b, T->ev[c][r][2], T->ev[c][r][3]));
                ms_out[0] += a;
                ms_out[1] += b;
            }
            ++timed;
        }
        if (timed > 0) {
            ms_out[0] /= static_cast<float>(timed);
            ms_out[1] /= static_cast<float>(timed);
        }
        if (T->made && T->calls > 0) {
            unsigned int n = 0;
            MOSAIC_HIP_CHECK(hipMemcpy(&n, T->fallbacks, sizeof(n), hipMemcpyDeviceToHost));
            ms_out[2] = static_cast<float>(n);
        }
    }
    T->on = enable != 0;
    return MOSAIC_OK;
}

extern "C" int mosaic_kin_sum(int64_t atom_count, int n_species, const int64_t* counts,
                              const mosaic_species_t* species_tab, const mosaic_kin_opts_t* opts,
                              int n_frames, const float* rot, const float* trans, int Ny, int Nz,
                              const double* coords, void* out, int out_is_c128,
                              const double* stage_stats, const double* pixel_geom,
                              void* workspace, size_t workspace_bytes, int* path_used,
                              mosaic_stream_t stream) {
    size_t need = 0;
    int rc = mosaic_kin_workspace(atom_count, n_species, counts, n_frames, Ny, Nz, &need);
    if (rc != MOSAIC_OK) return rc;
    if (!workspace || workspace_bytes < need) return MOSAIC_ERR_WORKSPACE;
    if (!species_tab || !opts || !rot || !trans || !coords || !out)
        return MOSAIC_ERR_ARGS;
    if (!(opts->k > 0.0) || (stage_stats && !(stage_stats[0] >= 0.0)) || opts->path < 0 ||
        opts->path > 7 || opts->analyser_kind < 0 || opts->analyser_kind > 2)
        return MOSAIC_ERR_ARGS;
    hipStream_t st = static_cast<hipStream_t>(stream);
    Layout L;
    rc = make_layout(workspace, atom_count, n_species, counts, n_frames, Ny, Nz, L);
    if (rc != MOSAIC_OK) return rc;
    const int npix = Ny * Nz;
    const double K = opts->k * INV_2PI_HI;
    if (path_used) *path_used = 0;
    if (atom_count == 0 || L.n_tiles == 0) return MOSAIC_OK;
    // statistics the caller left in the workspace: one read, one wait
    double ss_ws[2], pg_ws[N_GEOM];
    if (!stage_stats || !pixel_geom) {
        rc = read_stats(L, stage_stats ? nullptr : ss_ws, pixel_geom ? nullptr : pg_ws, st);
        if (rc != MOSAIC_OK) return rc;
        if (!stage_stats) stage_stats = ss_ws;
        if (!pixel_geom) pixel_geom = pg_ws;
        if (!(stage_stats[0] >= 0.0)) return MOSAIC_ERR_ARGS;
    }

    // Error bounds.  Per-tile path: tile radius against the nearest pixel
    // distance (tiles move rigidly with the sample pose).  Where it fails,
    // the automatic choice takes the exact per-pair kernel, which has no
    // such bound.
    const double radius_wl = stage_stats[0];
    const double rho_min = pixel_geom[0], du1 = pixel_geom[1], du2 = pixel_geom[2];
    const double dpatch = pixel_geom[3];
    const double rad_m = radius_wl / K;
    const double third = K * rad_m * rad_m * rad_m / (rho_min * rho_min);
    const bool tile_ok = rho_min > 0.0 && rad_m / rho_min <= 1e-6 && third <= 1e-7;
    const double second = tile_ok ? radius_wl * radius_wl / (2.0 * K * rho_min) : 0.0;
    const bool use_second = second > 1e-7;

    std::vector<FrameRec> fr(n_frames);
    double tmax = 0.0;
    for (int f = 0; f < n_frames; ++f) {
        for (int q = 0; q < 9; ++q) fr[f].R[q] = static_cast<double>(rot[9 * f + q]);
        for (int d = 0; d < 3; ++d) fr[f].T[d] = static_cast<double>(trans[3 * f + d]) / 1e10;
        for (int d = 0; d < 3; ++d) fr[f].xs[d] = fr[f].R[d];
        fr[f].pad = 0.0;
        tmax = std::max(tmax, std::sqrt(fr[f].T[0] * fr[f].T[0] + fr[f].T[1] * fr[f].T[1] +
                                        fr[f].T[2] * fr[f].T[2]));
    }
    // Recurrence path: every tile origin within 1e-6 rad of the lab origin as
    // seen from the nearest pixel (per-pixel form factor), the dropped third
    // order of the origin phase below 1e-8 rev, the row line fit below 1e-6
    // rev and the step phasor inside its Taylor range.
    const double omax = stage_stats[1] + tmax;
    const bool far = tile_ok && omax / rho_min <= 1e-6 &&
                     K * omax * omax * omax / (rho_min * rho_min) <= 1e-8;
    const bool rec_geom = far && 6.2831853 * du1 * radius_wl <= 0.05;
    int kp = 0;
    if (rec_geom && rec_curv_bound(8, du2, radius_wl) <= 1e-6) kp = 8;
    else if (rec_geom && rec_curv_bound(4, du2, radius_wl) <= 1e-6) kp = 4;
    // matrix-core sum: the patch model's deviation from the pixel grid,
    // times the largest offset, below 1e-6 rev
    // and the patch-constant second-order origin term: its change across a
    // patch, K |O|^2 |du| / rho, below 1e-7 rev
    // (and the atom phase |v.sigma| + 1 <= 2 radius + 1 inside the +-256
    // revolutions that v_sin / v_cos reduce exactly)
    const bool mma_ok = far && dpatch * radius_wl <= 1e-6 &&
                        K * omax * omax * (MMA_P * du1) / rho_min <= 1e-7 &&
                        2.0 * radius_wl + 2.0 <= 256.0;
    // moment sum: the same bounds with the deviation from the model that
    // includes the cross term
    auto poly_patch_ok = [&](double r) {
        return far && pixel_geom[6] * r <= 1e-6 &&
               K * omax * omax * (MMA_P * du1) / rho_min <= 1e-7 && 2.0 * r + 2.0 <= 256.0;
    };
    const int poly_d = poly_degree(poly_patch_ok(radius_wl), radius_wl, pixel_geom, K * omax, true);
    // The analyser weighs each ray (atom -> pixel) by its own direction: only
    // the exact kernel applies it.
    const int ana = opts->analyser_kind;
    int path = opts->path;
    // The probe form takes a few pixels (its lanes are the pixel-poses) in the
    // far field ahead of the per-tile form, which gives each lane a run of
    // pixels and starves on them.
    const bool probe_small = far && static_cast<int64_t>(npix) * n_frames <= PROBE_AUTO_OUTPUTS;
    // matrix-core moment sum: the moment sum's geometry bounds, the range of
    // the origin phase, and the expansion about each group's mean gradients;
    // the automatic choice takes it from one full region of 8 x 8 patches up
    // (with fewer, most of a region's waves would only build the basis)
    const int n_patches = ((Ny + MMA_P - 1) / MMA_P) * ((Nz + MMA_P - 1) / MMA_P);
    auto gm_ok = [&](double r) {
        return poly_patch_ok(r) && poly_degree(true, r, pixel_geom, K * omax, true) != 0 &&
               gm_bounds(r, pixel_geom);
    };
    const bool gm_path = n_patches >= GM_REG * GM_REG && gm_ok(radius_wl);
    if (path == 0)
        path = ana ? 5
                   : (gm_path ? 7
                              : (poly_d ? 4
                                        : (mma_ok ? 3
                                                  : (kp ? 2 : (probe_small ? 6 : (tile_ok ? 1 : 5))))));
    if ((path == 1 && !tile_ok) || (path == 2 && kp == 0) || (path == 3 && !mma_ok) ||
        (path == 4 && poly_d == 0) || (path == 6 && !far) || (path == 7 && !gm_ok(radius_wl)) ||
        (ana && path != 5))
        return MOSAIC_ERR_UNSUPPORTED;
    // Moment sum over units of 2 or 4 tiles: the largest unit radius for
    // which every radius-dependent bound above still holds at the chosen
    // degree and second-order setting.
    float grp_rmax = 0.f;
    if (path == 4 || path == 7) {
        auto holds = [&](double r) {
            const double rm = r / K;
            if (rm / rho_min > 1e-6 || K * rm * rm * rm / (rho_min * rho_min) > 1e-7) return false;
            if (!use_second && r * r / (2.0 * K * rho_min) > 1e-7) return false;
            if (path == 7) return gm_ok(r);
            const int d = poly_degree(poly_patch_ok(r), r, pixel_geom, K * omax, true);
            return d != 0 && d <= poly_d;
        };
        double lo = radius_wl, hi = 256.0;
        if (holds(lo)) {
            for (int it = 0; it < 60; ++it) {
                const double mid = 0.5 * (lo + hi);
                if (holds(mid)) lo = mid;
                else hi = mid;
            }
            grp_rmax = static_cast<float>(lo * (1.0 - 1e-6));
        }
    }

    MOSAIC_HIP_CHECK(hipMemcpyAsync(L.frames, fr.data(), sizeof(FrameRec) * n_frames,
                                    hipMemcpyHostToDevice, st));
    hipError_t e;
    if (path == 5) {
        ExactArgs A;
        A.atoms = L.atoms;
        A.tiles = L.tiles;
        A.frames = L.frames;
        A.cx = coords;
        A.cy = coords + npix;
        A.cz = coords + 2 * (int64_t)npix;
        A.theta0 = L.px.theta0;
        A.rho0 = L.px.rho0;
        A.qcut = L.px.qcut;
        A.partial = L.partial;
        A.n_tiles = L.n_tiles;
        A.tiles_per_slice = L.tiles_per_slice;
        A.n_slices = L.n_slices;
        A.npix = npix;
        A.K = K;
        A.k = opts->k;
        double an = 0.0;
        for (int d = 0; d < 3; ++d) {
            A.pol_e[d] = opts->pol_e[d];
            an += opts->analyser_axis[d] * opts->analyser_axis[d];
        }
        an = an > 0.0 ? 1.0 / std::sqrt(an) : 0.0;
        for (int d = 0; d < 3; ++d) A.ana_c[d] = opts->analyser_axis[d] * an;
        A.pol_deg = static_cast<float>(opts->pol_deg);
        A.ana_accept = static_cast<float>(opts->analyser_accept_rad);
        A.ana_inv_hw = opts->analyser_darwin_hw_rad > 0.0
                           ? static_cast<float>(1.0 / opts->analyser_darwin_hw_rad)
                           : 0.0f;
        A.remove_forward = opts->remove_forward ? 1 : 0;
        A.apply_decay = opts->apply_decay ? 1 : 0;
        for (int sp = 0; sp < n_species; ++sp)
            A.form[sp] = make_form_row(species_tab[sp].wk, species_tab[sp].f0_zero,
                                       species_tab[sp].fp, species_tab[sp].fpp, opts->k);
        const bool pol = opts->apply_pol != 0;
        if (ana == 1)
            e = pol ? launch_exact<true, 1>(A, L.n_slices, n_frames, st)
                    : launch_exact<false, 1>(A, L.n_slices, n_frames, st);
        else if (ana == 2)
            e = pol ? launch_exact<true, 2>(A, L.n_slices, n_frames, st)
                    : launch_exact<false, 2>(A, L.n_slices, n_frames, st);
        else
            e = pol ? launch_exact<true, 0>(A, L.n_slices, n_frames, st)
                    : launch_exact<false, 0>(A, L.n_slices, n_frames, st);
    } else if (path == 3 || path == 4) {
        MmaArgs A;
        A.atoms = L.atoms;
        A.tiles = L.tiles;
        A.frames = L.frames;
        A.u0h = L.px.u0h;
        A.theta0 = L.px.theta0;
        A.dup = L.px.dup;
        A.gmma = L.px.gmma;
        A.partial = L.partial;
        A.n_tiles = L.n_tiles;
        A.tiles_per_slice = L.tiles_per_slice;
        A.n_slices = L.n_slices;
        A.Ny = Ny;
        A.Nz = Nz;
        A.npix = npix;
        A.pby = (Ny + MMA_P - 1) / MMA_P;
        A.n_patches = A.pby * ((Nz + MMA_P - 1) / MMA_P);
        A.K = K;
        A.grp_r = L.grp_r;
        A.grp_rmax = grp_rmax;
        if (path == 3)
            e = use_second ? launch_mma<true>(A, L.n_slices, n_frames, st)
                           : launch_mma<false>(A, L.n_slices, n_frames, st);
        else if (poly_d == 5)
            e = use_second ? launch_poly<5, true>(A, L.n_slices, n_frames, st)
                           : launch_poly<5, false>(A, L.n_slices, n_frames, st);
        else
            e = use_second ? launch_poly<6, true>(A, L.n_slices, n_frames, st)
                           : launch_poly<6, false>(A, L.n_slices, n_frames, st);
    } else if (path == 7) {
        GmArgs A;
        A.atoms = L.atoms;
        A.tiles = L.tiles;
        A.frames = L.frames;
        A.u0h = L.px.u0h;
        A.theta0 = L.px.theta0;
        A.dup = L.px.dup;
        A.gmma = L.px.gmma;
        A.partial = L.partial;
        A.mom = L.gm_mom;
        A.heads = L.gm_head;
        A.grp_r = L.grp_r;
        A.grp_rmax = grp_rmax;
        A.n_tiles = L.n_tiles;
        A.tiles_per_slice = L.tiles_per_slice;
        A.n_slices = L.n_slices;
        A.n_frames = n_frames;
        A.Ny = Ny;
        A.Nz = Nz;
        A.npix = npix;
        A.pby = (Ny + MMA_P - 1) / MMA_P;
        A.pbz = (Nz + MMA_P - 1) / MMA_P;
        A.n_patches = n_patches;
        A.ngy = (A.pby + GM_REG - 1) / GM_REG;
        A.n_groups = A.ngy * ((A.pbz + GM_REG - 1) / GM_REG);
        A.K = K;
        // rounds of unit heads whose moments fit the buffer
        const int64_t per_tile = static_cast<int64_t>(n_patches) * n_frames * GM_REC;
        const int tiles_per_round = static_cast<int>(
            std::max<int64_t>(1, std::min<int64_t>(L.n_tiles, static_cast<int64_t>(L.gm_count) / per_tile)));
        // moments kernel: about 4096 waves over the round's tiles
        const int64_t waves = static_cast<int64_t>(A.n_groups) * GM_WAVES * n_frames;
        KinTiming* T = timing_state();
        const int slot = timing_slot(T);
        A.fallbacks = slot >= 0 ? T->fallbacks : nullptr;
        A.pixel_eval = opts->gm_pixel_eval ? 1 : 0;
        hipLaunchKernelGGL(k_gm_heads, dim3(blocks_for(L.n_tiles, 256)), dim3(256), 0, st,
                           L.grp_r, grp_rmax, L.n_tiles, L.gm_head);
        e = hipGetLastError();
        int round = 0;
        for (int tb = 0; tb < L.n_tiles && e == hipSuccess; tb += tiles_per_round, ++round) {
            const int te = std::min(L.n_tiles, tb + tiles_per_round);
            const int rt = te - tb;
            const int want = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(rt, (4096 + waves - 1) / waves)));
            A.a_tiles_per_slice = (rt + want - 1) / want;
            const int slices_a = (rt + A.a_tiles_per_slice - 1) / A.a_tiles_per_slice;
            A.t_begin = tb;
            A.t_end = te;
            A.accumulate = tb > 0 ? 1 : 0;
            timing_mark(T, slot, round, 0, st);
            e = use_second ? launch_gm_moments<true>(A, slices_a, st)
                           : launch_gm_moments<false>(A, slices_a, st);
            timing_mark(T, slot, round, 1, st);
            if (e != hipSuccess) break;
            timing_mark(T, slot, round, 2, st);
            hipLaunchKernelGGL(k_gm_eval, dim3((n_patches + POLY_WAVES - 1) / POLY_WAVES, L.n_slices, n_frames),
                               dim3(64 * POLY_WAVES), 0, st, A);
            timing_mark(T, slot, round, 3, st);
            e = hipGetLastError();
        }
        timing_done(T, slot, round);
    } else if (path == 6) {
        ProbeArgs A;
        A.atoms = L.atoms;
        A.tiles = L.tiles;
        A.frames = L.frames;
        A.u0h = L.px.u0h;
        A.theta0 = L.px.theta0;
        A.gtab = L.px.gtab;
        A.partial = L.partial;
        A.n_tiles = L.n_tiles;
        A.tiles_per_slice = L.tiles_per_slice;
        A.n_slices = L.n_slices;
        A.npix = npix;
        A.n_frames = n_frames;
        A.K = K;
        e = use_second ? launch_probe<true>(A, L.n_slices, st)
                       : launch_probe<false>(A, L.n_slices, st);
    } else if (path == 2) {
        RecArgs A;
        A.atoms = L.atoms;
        A.tiles = L.tiles;
        A.frames = L.frames;
        A.u0h = L.px.u0h;
        A.theta0 = L.px.theta0;
        A.gtab = L.px.gtab;
        A.partial = L.partial;
        A.n_tiles = L.n_tiles;
        A.tiles_per_slice = L.tiles_per_slice;
        A.n_slices = L.n_slices;
        A.Ny = Ny;
        A.Nz = Nz;
        A.npix = npix;
        A.nbx = (Ny + kp - 1) / kp;
        A.K = K;
        if (kp == 8)
            e = use_second ? launch_rec<8, true>(A, L.n_slices, n_frames, st)
                           : launch_rec<8, false>(A, L.n_slices, n_frames, st);
        else
            e = use_second ? launch_rec<4, true>(A, L.n_slices, n_frames, st)
                           : launch_rec<4, false>(A, L.n_slices, n_frames, st);
    } else {
        SumArgs A;
        A.atoms = L.atoms;
        A.tiles = L.tiles;
        A.frames = L.frames;
        A.cx = coords;
        A.cy = coords + npix;
        A.cz = coords + 2 * (int64_t)npix;
        A.theta0 = L.px.theta0;
        A.rho0 = L.px.rho0;
        A.qcut = L.px.qcut;
        A.partial = L.partial;
        A.n_tiles = L.n_tiles;
        A.tiles_per_slice = L.tiles_per_slice;
        A.n_slices = L.n_slices;
        A.npix = npix;
        A.K = K;
        A.k = opts->k;
        A.pol_deg = opts->pol_deg;
        for (int d = 0; d < 3; ++d) A.pol_e[d] = opts->pol_e[d];
        A.remove_forward = opts->remove_forward ? 1 : 0;
        A.apply_pol = opts->apply_pol ? 1 : 0;
        A.apply_decay = opts->apply_decay ? 1 : 0;
        A.n_species = n_species;
        fill_species(species_tab, n_species, A.spc);
        const int ppt = pix_per_thread(npix);
        if (ppt == 8)
            e = use_second ? launch_sum<8, true>(A, L.pix_blocks, L.n_slices, n_frames, st)
                           : launch_sum<8, false>(A, L.pix_blocks, L.n_slices, n_frames, st);
        else
            e = use_second ? launch_sum<4, true>(A, L.pix_blocks, L.n_slices, n_frames, st)
                           : launch_sum<4, false>(A, L.pix_blocks, L.n_slices, n_frames, st);
    }
    if (e != hipSuccess) return MOSAIC_ERR_HIP;
    const int64_t nq = static_cast<int64_t>(npix) * n_frames;
    if (out_is_c128)
        hipLaunchKernelGGL(k_reduce<true>, dim3(blocks_for(nq, 256)), dim3(256), 0, st, L.partial,
                           L.n_slices, npix, n_frames, out);
    else
        hipLaunchKernelGGL(k_reduce<false>, dim3(blocks_for(nq, 256)), dim3(256), 0, st,
                           L.partial, L.n_slices, npix, n_frames, out);
    MOSAIC_HIP_CHECK(hipGetLastError());
    if (path_used)
        *path_used = path == 5   ? 50 + ana
                     : path == 6 ? 60 + (use_second ? 100 : 0)
                     : path == 7 ? 76 + (use_second ? 100 : 0)
                     : path == 4 ? 40 + poly_d + (use_second ? 100 : 0)
                     : path == 3 ? 32 + (use_second ? 100 : 0)
                     : path == 2 ? (kp == 8 ? 28 : 24) + (use_second ? 100 : 0)
                                 : (use_second ? 12 : 11);
    return MOSAIC_OK;
}

extern "C" int mosaic_ein_bilinear(int64_t atom_count, const float* pos, const float* rot,
                                   const float* trans, const float* E0, int NyB, int NzB,
                                   double du, double dv, double uc, double vc, float* amp_out,
                                   mosaic_stream_t stream) {
    if (atom_count < 0 || !rot || !trans || NyB < 1 || NzB < 1) return MOSAIC_ERR_ARGS;
    if (atom_count == 0) return MOSAIC_OK;
    if (!pos || !E0 || !amp_out || !(du > 0.0) || !(dv > 0.0)) return MOSAIC_ERR_ARGS;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_ein, dim3(blocks_for(atom_count, 256)), dim3(256), 0, st, pos,
                       atom_count, rot[0], rot[1], rot[2], rot[3], rot[4], rot[5], rot[6], rot[7],
                       rot[8], trans[0], trans[1], trans[2],
                       reinterpret_cast<const float2*>(E0), NyB, NzB, du, dv, uc, vc,
                       reinterpret_cast<float2*>(amp_out));
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

// Pixel moments of the moment adjoint's units: per tile (unit heads use
// theirs) and (frame, patch) pair, 2 adj_moments(D) floats at the largest D.
static size_t adjoint_scratch_bytes(int n_tiles, int n_frames, int Ny, int Nz) {
    const size_t patches = static_cast<size_t>((Ny + MMA_P - 1) / MMA_P) * ((Nz + MMA_P - 1) / MMA_P);
    constexpr size_t NMF = 2 * adj_moments(ADJ_MAX_DEGREE);
    return static_cast<size_t>(std::max(1, n_tiles)) * n_frames * patches * NMF * sizeof(float);
}

extern "C" int mosaic_kin_adjoint_scratch(int64_t atom_count, int n_species, const int64_t* counts,
                                          int n_frames, int Ny, int Nz, size_t* bytes) {
    size_t need = 0;
    const int rc = mosaic_kin_workspace(atom_count, n_species, counts, n_frames, Ny, Nz, &need);
    if (rc != MOSAIC_OK) return rc;
    *bytes = adjoint_scratch_bytes(count_tiles(n_species, counts), n_frames, Ny, Nz);
    return MOSAIC_OK;
}

extern "C" int mosaic_kin_adjoint(int64_t atom_count, int n_species, const int64_t* counts,
                                  double k, int n_frames, const float* rot, const float* trans,
                                  int Ny, int Nz, const double* stage_stats,
                                  const double* pixel_geom,
                                  const float* grad_field, float* grad_pos, void* workspace,
                                  size_t workspace_bytes, void* scratch, size_t scratch_bytes,
                                  int path, mosaic_stream_t stream) {
    size_t need = 0;
    int rc = mosaic_kin_workspace(atom_count, n_species, counts, n_frames, Ny, Nz, &need);
    if (rc != MOSAIC_OK) return rc;
    if (!workspace || workspace_bytes < need) return MOSAIC_ERR_WORKSPACE;
    if (!rot || !trans || !stage_stats || !pixel_geom || !grad_field || !grad_pos)
        return MOSAIC_ERR_ARGS;
    if (path != 0 && path != 2 && path != 3 && path != 4) return MOSAIC_ERR_ARGS;
    hipStream_t st = static_cast<hipStream_t>(stream);
    Layout L;
    rc = make_layout(workspace, atom_count, n_species, counts, n_frames, Ny, Nz, L);
    if (rc != MOSAIC_OK) return rc;
    if (atom_count == 0 || L.n_tiles == 0) return MOSAIC_OK;
    // The adjoint reads the staged sample and the pixel tables of the forward
    // and uses the row line model: it declines where the forward recurrence
    // would.
    const double radius_wl = stage_stats[0];
    const double rho_min = pixel_geom[0], du1 = pixel_geom[1], du2 = pixel_geom[2];
    std::vector<FrameRec> fr(n_frames);
    double tmax = 0.0;
    for (int f = 0; f < n_frames; ++f) {
        for (int q = 0; q < 9; ++q) fr[f].R[q] = static_cast<double>(rot[9 * f + q]);
        for (int d = 0; d < 3; ++d) fr[f].T[d] = static_cast<double>(trans[3 * f + d]) / 1e10;
        for (int d = 0; d < 3; ++d) fr[f].xs[d] = fr[f].R[d];
        fr[f].pad = 0.0;
        tmax = std::max(tmax, std::sqrt(fr[f].T[0] * fr[f].T[0] + fr[f].T[1] * fr[f].T[1] +
                                        fr[f].T[2] * fr[f].T[2]));
    }
    const double Kk = k * INV_2PI_HI;
    if (!(Kk > 0.0) || !(rho_min > 0.0)) return MOSAIC_ERR_ARGS;
    const double rad_m = radius_wl / Kk;
    if (rad_m / rho_min > 1e-6 || Kk * rad_m * rad_m * rad_m / (rho_min * rho_min) > 1e-7)
        return MOSAIC_ERR_UNSUPPORTED;
    const double second = radius_wl * radius_wl / (2.0 * Kk * rho_min);
    const bool use_second = second > 1e-7;
    const double omax = stage_stats[1] + tmax;
    const bool far = omax / rho_min <= 1e-6 &&
                     Kk * omax * omax * omax / (rho_min * rho_min) <= 1e-8;
    const bool ok = far && 6.2831853 * du1 * radius_wl <= 0.05;
    // the matrix-core form under the forward's bounds, else the recurrence
    // (the atom phases, |v.sigma| + 1 <= 2 radius + 1 revolutions, go to
    // v_sin / v_cos unreduced, so they must stay inside their +-256 range; the
    // moment form takes the same gate)
    const bool mma_ok = far && pixel_geom[3] * radius_wl <= 1e-6 &&
                        Kk * omax * omax * (MMA_P * du1) / rho_min <= 1e-7 &&
                        2.0 * radius_wl + 2.0 <= 256.0;
    int kp = 0;
    if (ok && rec_curv_bound(16, du2, radius_wl) <= 1e-6) kp = 16;
    else if (ok && rec_curv_bound(8, du2, radius_wl) <= 1e-6) kp = 8;
    // the moment form under the forward's truncation bounds
    const int poly_d = poly_degree(mma_ok, radius_wl, pixel_geom, Kk * omax);
    const bool use_poly = path == 4 || (path == 0 && poly_d > 0);
    const bool use_mma = !use_poly && (path == 3 || (path == 0 && mma_ok));
    if ((use_poly && poly_d == 0) || (use_mma && !mma_ok) || (!use_poly && !use_mma && kp == 0))
        return MOSAIC_ERR_UNSUPPORTED;
    MOSAIC_HIP_CHECK(hipMemcpyAsync(L.frames, fr.data(), sizeof(FrameRec) * n_frames,
                                    hipMemcpyHostToDevice, st));
    const int64_t n_gin0 = 2 * static_cast<int64_t>(n_frames) * Ny * Nz;
    const int64_t n_gtab0 = 2 * static_cast<int64_t>(n_species) * Ny * Nz;
    if (use_mma || use_poly) {
        AdjMmaArgs M;
        M.atoms = L.atoms;
        M.tiles = L.tiles;
        M.frames = L.frames;
        M.u0h = L.px.u0h;
        M.theta0 = L.px.theta0;
        M.dup = L.px.dup;
        M.gmma = L.px.gmma;
        M.gin = reinterpret_cast<const float2*>(grad_field);
        M.amax = L.amax_blocks;
        M.grad = grad_pos;
        M.n_frames = n_frames;
        M.Ny = Ny;
        M.Nz = Nz;
        M.npix = Ny * Nz;
        M.pby = (Ny + MMA_P - 1) / MMA_P;
        M.n_patches = M.pby * ((Nz + MMA_P - 1) / MMA_P);
        M.K = Kk;
        hipLaunchKernelGGL(k_amax, dim3(AMAX_BLOCKS), dim3(256), 0, st, grad_field, n_gin0,
                           reinterpret_cast<const float*>(L.px.gmma), n_gtab0, L.amax_blocks);
        // the column curvature of the gradient direction, |hz| b^2 with
        // hz ~ du2z / 2 and |b| <= MMA_P / 2, may be left out below 1e-6
        const bool flat = 0.125 * MMA_P * MMA_P * pixel_geom[4] <= 1e-6;
        const dim3 grid(L.n_tiles), blk(64 * MMA_WAVES);
        const dim3 pblk(64 * POLY_WAVES);
        // units of tiles with a scratch for their pixel moments: the largest
        // unit radius for which the bounds above hold at the chosen degree
        float grp_rmax = 0.f;
        if (use_poly && scratch &&
            scratch_bytes >= adjoint_scratch_bytes(L.n_tiles, n_frames, Ny, Nz)) {
            auto holds = [&](double r) {
                const double rm = r / Kk;
                if (rm / rho_min > 1e-6 || Kk * rm * rm * rm / (rho_min * rho_min) > 1e-7)
                    return false;
                if (!use_second && r * r / (2.0 * Kk * rho_min) > 1e-7) return false;
                const bool ok_r = far && pixel_geom[3] * r <= 1e-6 &&
                                  Kk * omax * omax * (MMA_P * du1) / rho_min <= 1e-7 &&
                                  2.0 * r + 2.0 <= 256.0;
                const int d = poly_degree(ok_r, r, pixel_geom, Kk * omax);
                return d != 0 && d <= poly_d;
            };
            double lo = radius_wl, hi = 256.0;
            if (holds(lo)) {
                for (int it = 0; it < 60; ++it) {
                    const double mid = 0.5 * (lo + hi);
                    if (holds(mid)) lo = mid;
                    else hi = mid;
                }
                grp_rmax = static_cast<float>(lo * (1.0 - 1e-6));
            }
        }
        M.grp_r = L.grp_r;
        M.grp_rmax = grp_rmax;
        M.n_tiles = L.n_tiles;
        M.mom = static_cast<float*>(scratch);
        if (use_poly && grp_rmax > 0.f) {
            const int npairs = n_frames * M.n_patches;
            const dim3 mgrid(L.n_tiles, (npairs + ADJ_GROUP - 1) / ADJ_GROUP);
            if (poly_d == 5) hipLaunchKernelGGL((k_adjoint_moments<5>), mgrid, pblk, 0, st, M);
            else hipLaunchKernelGGL((k_adjoint_moments<6>), mgrid, pblk, 0, st, M);
            MOSAIC_HIP_CHECK(hipGetLastError());
            if (poly_d == 5 && use_second)
                hipLaunchKernelGGL((k_adjoint_poly<5, true, true>), grid, pblk, 0, st, M);
            else if (poly_d == 5)
                hipLaunchKernelGGL((k_adjoint_poly<5, false, true>), grid, pblk, 0, st, M);
            else if (use_second)
                hipLaunchKernelGGL((k_adjoint_poly<6, true, true>), grid, pblk, 0, st, M);
            else
                hipLaunchKernelGGL((k_adjoint_poly<6, false, true>), grid, pblk, 0, st, M);
        } else if (use_poly) {
            if (poly_d == 5 && use_second)
                hipLaunchKernelGGL((k_adjoint_poly<5, true, false>), grid, pblk, 0, st, M);
            else if (poly_d == 5)
                hipLaunchKernelGGL((k_adjoint_poly<5, false, false>), grid, pblk, 0, st, M);
            else if (use_second)
                hipLaunchKernelGGL((k_adjoint_poly<6, true, false>), grid, pblk, 0, st, M);
            else
                hipLaunchKernelGGL((k_adjoint_poly<6, false, false>), grid, pblk, 0, st, M);
        } else if (use_second) {
            if (flat) hipLaunchKernelGGL((k_adjoint_mma<true, 2>), grid, blk, 0, st, M);
            else hipLaunchKernelGGL((k_adjoint_mma<true, 3>), grid, blk, 0, st, M);
        } else {
            if (flat) hipLaunchKernelGGL((k_adjoint_mma<false, 2>), grid, blk, 0, st, M);
            else hipLaunchKernelGGL((k_adjoint_mma<false, 3>), grid, blk, 0, st, M);
        }
        MOSAIC_HIP_CHECK(hipGetLastError());
        return MOSAIC_OK;
    }
    AdjArgs A;
    A.atoms = L.atoms;
    A.tiles = L.tiles;
    A.frames = L.frames;
    A.u0h = L.px.u0h;
    A.theta0 = L.px.theta0;
    A.gtab = L.px.gtab;
    A.gin = reinterpret_cast<const float2*>(grad_field);
    A.amax = L.amax_blocks;
    A.grad = grad_pos;
    A.n_frames = n_frames;
    A.Ny = Ny;
    A.Nz = Nz;
    A.npix = Ny * Nz;
    A.nbx = (Ny + kp - 1) / kp;
    A.K = Kk;
    const int64_t n_gin = 2 * static_cast<int64_t>(n_frames) * Ny * Nz;
    const int64_t n_gtab = 2 * static_cast<int64_t>(n_species) * Ny * Nz;
    hipLaunchKernelGGL(k_amax, dim3(AMAX_BLOCKS), dim3(256), 0, st, grad_field, n_gin,
                       reinterpret_cast<const float*>(L.px.gtab), n_gtab, L.amax_blocks);
    const dim3 grid(L.n_tiles), blk(SUM_THREADS);
    if (kp == 16) {
        if (use_second) hipLaunchKernelGGL((k_adjoint<16, true>), grid, blk, 0, st, A);
        else hipLaunchKernelGGL((k_adjoint<16, false>), grid, blk, 0, st, A);
    } else {
        if (use_second) hipLaunchKernelGGL((k_adjoint<8, true>), grid, blk, 0, st, A);
        else hipLaunchKernelGGL((k_adjoint<8, false>), grid, blk, 0, st, A);
    }
    MOSAIC_HIP_CHECK(hipGetLastError());
    return MOSAIC_OK;
}

