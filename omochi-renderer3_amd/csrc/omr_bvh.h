// omr_bvh.h -- host-side SAH builder that flattens to 4-wide and 8-wide nodes.
//
// Replaces the reference's two-stage build (BVH::Construct with full per-axis
// sorts, src/renderer/BVH.cpp:90-312, collapsed by QBVH::Construct_internal,
// src/renderer/QBVH.cpp:164-297) with a binned SAH build that runs in
// parallel over subtrees and collapses by repeatedly opening the child with
// the largest surface area.  One binary tree is collapsed twice, to 4-wide
// and to 8-wide nodes whose leaves are the same primitive runs, so both share
// the primitive records and the traversal picks either (the node width is a
// launch tuning).  The tree shape is not a parity target; the closest-hit
// result is (the box test is conservative).
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "omr_layout.h"

namespace omr {

struct BuildInput {
  size_t n_tri = 0;
  const float *tri = nullptr;  // n_tri x 9: v0, e1, e2
  size_t n_sph = 0;
  const float *sph = nullptr;  // n_sph x 4
};

struct BuildOutput {
  std::vector<WNode4> nodes;
  std::vector<WNode8> nodes8;       // the same leaves under 8-wide nodes
  std::vector<PrimRec> prims;       // leaf order (of the 4-wide collapse)
  std::vector<uint32_t> prim_order; // leaf slot -> original primitive id
  uint32_t depth = 0;               // levels of 4-wide nodes
  uint32_t stack_bound = 0;         // max traversal stack entries needed
  uint32_t depth8 = 0;              // levels of 8-wide nodes
  uint32_t stack_bound8 = 0;
  float pad = 0.0f;                 // absolute box padding used
  double build_ms = 0.0;
};

// Returns false (with a message) on invalid input (non-finite coordinates,
// negative radius, too many primitives).
bool build_wide_bvh(const BuildInput &in, BuildOutput &out, std::string &err, int threads = 0);

}  // namespace omr
