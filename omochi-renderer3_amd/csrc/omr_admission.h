// omr_admission.h -- host-side bookkeeping of served mode: the request arena,
// the order and memory rules of admission, and the latency record.  Pure host
// code with no device calls; the server loop (omr_server.cpp) uses it on the
// wall clock, and omr_serve_simulate drives the same code with a fake clock
// and a fixed-rate device model (deterministic tests, capacity planning).
#pragma once
#include <cstddef>
#include <cstdint>
#include <deque>
#include <map>
#include <vector>

#include "../../include/omr_capi.h"

namespace omr {

inline uint64_t align256(uint64_t b) { return (b + 255) & ~uint64_t(255); }

inline uint64_t samples_of(const omr_camera &c, const omr_render_params &p) {
  return (uint64_t)c.width * c.height * p.supersamples * p.supersamples * (uint64_t)(p.spp_end - p.spp_begin);
}

// Device bytes of a request's fixed-point sums (int64 RGB) and resolved image
// (float RGB), each section 256-byte aligned in its block.
inline uint64_t image_bytes(uint64_t width, uint64_t height) {
  const uint64_t px = width * height;
  return align256(px * 3 * 8) + align256(px * 3 * 4);
}

// Per-sample records: 3 x int64 per sample.
inline uint64_t record_bytes(uint64_t samples) { return align256(samples * 24); }

// First-fit allocator over one block of `size` bytes in 256-byte granules:
// a request takes the lowest free range that holds it; a returned range
// merges with its free neighbours.
class Arena {
 public:
  void reset(uint64_t size);
  bool take(uint64_t bytes, uint64_t *off);
  void give(uint64_t off, uint64_t bytes);
  uint64_t size() const { return size_; }
  uint64_t free_bytes() const;

 private:
  std::map<uint64_t, uint64_t> free_;  // offset -> length
  uint64_t size_ = 0;
};

// What a request holds while it is live.
struct Placement {
  uint64_t image = 0;    // bytes of its sums and image
  uint64_t records = 0;  // bytes of its per-sample records; 0: pixel atomics (the same results)
  uint64_t off = 0;      // its block in the arena
  uint64_t bytes() const { return image + records; }
};

// At submission: the request's image and, when they fit the per-request
// record limit and (with the image) the whole arena, its records.  False when
// the image alone exceeds the arena: the request can never be admitted.
bool plan_request(uint64_t image, uint64_t samples, uint64_t record_limit, uint64_t arena_size, Placement *p);

// At admission: a block for the request, without its records when only that
// fits.  False when not even the image fits now.
bool place_request(Arena &arena, Placement *p);

// In-order admission.  Requests leave the front of `queue` while the live
// window has room and each gets a block; the first one that finds no block
// waits, and every request behind it with it, so requests are admitted in
// submission order and a large one is never passed over.  `skip(r)` is true
// for a request to drop without a block (cancelled); `drop(r)` receives it.
// `place(r)` returns the request's Placement; `admit(r)` receives each
// admitted request.  Returns the number admitted.
template <class R, class Skip, class Drop, class Place, class Admit>
size_t admit_in_order(std::deque<R> &queue, Arena &arena, size_t window_room, Skip skip, Drop drop, Place place,
                      Admit admit) {
  size_t n = 0;
  while (!queue.empty() && n < window_room) {
    R r = queue.front();
    if (skip(r)) {
      queue.pop_front();
      drop(r);
      continue;
    }
    if (!place_request(arena, place(r))) break;
    queue.pop_front();
    admit(r);
    n++;
  }
  return n;
}

// Latencies of completed requests and their percentiles: the value at rank
// round(q * (n - 1)) (halves up) of the sorted latencies.
class LatencyBook {
 public:
  void clear() { v_.clear(); }
  void add(double ms) { v_.push_back(ms); }
  size_t count() const { return v_.size(); }
  double percentile(double q) const;
  double max() const;

 private:
  std::vector<double> v_;
};

// Discrete-event run of the admission code against a fake clock; see
// omr_serve_simulate in include/omr_capi.h.
int serve_simulate(const omr_sim_request *req, size_t n, const omr_sim_config &cfg, omr_sim_outcome *out,
                   omr_sim_summary *summary);

}  // namespace omr
