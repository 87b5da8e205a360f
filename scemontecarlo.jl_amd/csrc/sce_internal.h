// sce_internal.h — shared internals of libsce_hip.so (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <vector>

#include "../../include/sce_hip.h"
#include "sce_math.h"

// Records the thread-local error message and returns `code`.
int32_t sce_fail(int32_t code, const char* fmt, ...);

// One flattened walk entry of a sublattice program: the weight, the three factor
// slots of the staged neighbour block (index nb*nlm + row; unused factors point
// at a slot holding 1.0) and the target row of the site's own tesseral row.
struct alignas(16) WalkEntry {
  double w;
  uint16_t f0, f1, f2, tgt;
};
static_assert(sizeof(WalkEntry) == 16, "WalkEntry must be 16 bytes");

// Per-sublattice device layout of a tiled Hamiltonian.
//
// For a TiledHamiltonian every site of training-cell atom a touches exactly the
// same site programs (one per (template, member slot) whose slot sits on atom
// a), with neighbour sites displaced by fixed cell offsets.  The walk of a site
// of sublattice a is therefore one shared list: adjacency entries in ascending
// program id, each expanded into its program entries, with factor columns
// addressed through a per-sublattice neighbour stencil (atom, cell offset).
struct HostLayout {
  int na = 0;
  int nb_max = 0;   // largest neighbour stencil
  int w_max = 0;    // largest walk (entries)
  std::vector<int32_t> walk_ptr;  // na + 1
  std::vector<WalkEntry> walk;
  std::vector<int32_t> nb_ptr;    // na + 1
  // per sublattice, the walk offset (relative to walk_ptr[a]) where each of its
  // pid-sorted adjacency entries' program entries start, plus the end
  std::vector<int32_t> prog_ptr;  // na + 1, into prog_off
  std::vector<int32_t> prog_off;
  std::vector<int32_t> nb;        // 4 ints per neighbour: atom (0-based), dc1, dc2, dc3
  // rotation-invariant surrogate descriptor of a site (per sublattice):
  // features [e.e_n for the NB stencil neighbours, (e.e_n)^2 for them, then one
  // per cluster spec]; spec (kind, x, y, z): kind 1 = e.(e_x x e_y) (triplet),
  // kind 2 = (e.e_x)(e_y.e_z) + (e.e_y)(e_x.e_z) + (e.e_z)(e_x.e_y) (body 4),
  // kind 3 = spin component x of e (body-1 terms, 3 specs)
  std::vector<int32_t> feat_ptr;  // na + 1, into feat (units of specs)
  std::vector<int32_t> feat;      // 4 ints per spec
  int nfeat_max = 0;              // largest 2*NB + specs
  // the representative site's pid-sorted adjacency, per sublattice (what every
  // site of the sublattice must repeat, translated): program id, path class
  // (1 pair, 2 triplet, 0 general) and stencil index of each factor
  std::vector<int32_t> rep_ptr;   // na + 1, into the rep_* arrays
  std::vector<int32_t> rep_pid;
  std::vector<int8_t> rep_cls;
  std::vector<int32_t> rep_nb;    // 4 ints per entry: factor count, stencil indices (-1 unused)
};

// Derives the layout from the reference tables (the representative site of each
// sublattice) and, when `verify`, checks on the host for every active site that
// its adjacency is the sublattice stencil translated to its cell.  Returns
// SCE_ELAYOUT (with a message) for tables that are not such a tiling.
int32_t sce_derive_layout(const sce_tables_desc* d, HostLayout* out, bool verify = true);

// The per-position tables a device build makes (the host restatement): for
// every colour position q, the nb_max neighbour sites (0-based, padded with the
// site itself); when nb_max <= 64 the lower-colour mask of its slots; the raw
// adjacency index -> pid-sorted index permutation (adj_max per position).
int32_t sce_position_tables_host(const sce_tables_desc* d, const HostLayout& L, std::vector<int32_t>& nbq,
                                 std::vector<unsigned long long>& low, std::vector<uint16_t>& perm, int* adj_max);
