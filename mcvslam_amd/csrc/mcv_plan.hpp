// Work descriptors shared by the host planner (mcv_orb_host.cpp) and the
// ORB kernels (mcv_orb.hip).  A plan covers a batch of images of arbitrary
// sizes: every pyramid level of every image is a LevelDesc, every FAST cell
// window is a CellDesc, every output keypoint slot belongs to one level.
#pragma once
#include <stdint.h>

namespace mcv {

constexpr int kMaxLevels = 16;
constexpr int kOctreeDepth = 15;        // quadtree depths encoded per candidate
constexpr int kSortCapLds = 4096;       // candidates per level sorted in LDS (more: pool memory)
constexpr int kHeapTabLds = 2048;       // node-table entries per level the octree heap replay stages in LDS
constexpr long long kOctLargePixels = 600000;  // levels this large get 1024-thread octree builds
constexpr int kMaxWin = 96;             // max FAST cell window side (pixels)
constexpr int kStripW = 256;            // strip kernels: one wave per 256 x 32 pixels,
constexpr int kStripH = 32;             //   4 adjacent pixels per lane
constexpr int kPitchAlign = 64;         // level rows start on 64-byte boundaries

struct LevelDesc {
    int img;            // image index in the batch
    int level;
    int w, h;           // level size
    int pitch;          // row stride in bytes: w rounded up to kPitchAlign
    int src;            // LevelDesc index of the previous level (-1 for level 0)
    int nfeat;          // octree target for this level (mnFeaturesPerLevel)
    long long pix;      // byte offset of the level in the pyramid / blur buffers
    int coef_x, coef_y; // resize tables: int2 (xofs, a0 | a1 << 16) per dst column / row
    int cell_first, cell_count;
    int minBX, maxBX, minBY, maxBY;  // ComputeKeyPointsOctTree borders
    int cand_base;      // first slot of the level's contiguous candidate array
    int cand_cap;       // capacity of that array
    int n_ini;          // octree root count
    float hX;           // octree root width
    int out_cap;        // keypoint slots reserved for the octree output
    int seed_first, seed_count;  // caller keypoints appended to this level
    int kp_slot;        // first per-level keypoint slot (octree output then seeds)
    float scale;        // mvScaleFactor[level]
    float patch;        // (float)(int)(31 * scale)
    int first_level;    // LevelDesc index of level 0 of the same image
    int nlevels;
};

struct CellDesc {
    int lvl;            // LevelDesc index
    int x0, y0, x1, y1; // FAST window [x0,x1) x [y0,y1) in level pixels
    int offx, offy;     // j*wCell, i*hCell
    int cap_base;       // candidate slots of this cell
};

// Octree state of one level after the build kernel.
struct OctLevel {
    int n;              // candidates
    int n_roots;        // depth-0 nodes
    int depth;          // deepest table depth (all nodes there are singletons)
    int failed;         // pool overflow or oversize level
    long long tab;      // u32 offset in the pool: start[T] then child[T]
    int total;          // T = number of node entries (with per-depth sentinels)
    long long sorted;   // u32 offset in the pool: candidate index per sorted position
    int nout;           // nodes emitted by the heap (octree output count)
    int pad;
};

}  // namespace mcv
