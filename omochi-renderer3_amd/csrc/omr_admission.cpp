// omr_admission.cpp -- host-side bookkeeping of served mode (see
// omr_admission.h) and its discrete-event run against a fake clock.
#include "omr_admission.h"

#include <algorithm>
#include <cmath>
#include <limits>

#include "omr_layout.h"

namespace omr {

void Arena::reset(uint64_t size) {
  size_ = size & ~uint64_t(255);
  free_.clear();
  if (size_) free_[0] = size_;
}

bool Arena::take(uint64_t bytes, uint64_t *off) {
  bytes = align256(bytes);
  if (bytes == 0) bytes = 256;
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < bytes) continue;
    *off = it->first;
    const uint64_t rest = it->second - bytes, at = it->first + bytes;
    free_.erase(it);
    if (rest) free_[at] = rest;
    return true;
  }
  return false;
}

void Arena::give(uint64_t off, uint64_t bytes) {
  bytes = align256(bytes);
  if (bytes == 0) bytes = 256;
  auto it = free_.emplace(off, bytes).first;
  auto next = std::next(it);
  if (next != free_.end() && it->first + it->second == next->first) {  // merge with the range after
    it->second += next->second;
    free_.erase(next);
  }
  if (it != free_.begin()) {  // merge with the range before
    auto prev = std::prev(it);
    if (prev->first + prev->second == it->first) {
      prev->second += it->second;
      free_.erase(it);
    }
  }
}

uint64_t Arena::free_bytes() const {
  uint64_t s = 0;
  for (const auto &f : free_) s += f.second;
  return s;
}

bool plan_request(uint64_t image, uint64_t samples, uint64_t record_limit, uint64_t arena_size, Placement *p) {
  p->image = image;
  p->off = 0;
  const uint64_t rec = record_bytes(samples);
  p->records = (rec <= record_limit && image + rec <= arena_size) ? rec : 0;
  return image <= arena_size;
}

bool place_request(Arena &arena, Placement *p) {
  if (arena.take(p->bytes(), &p->off)) return true;
  if (p->records == 0 || !arena.take(p->image, &p->off)) return false;
  p->records = 0;
  return true;
}

double LatencyBook::percentile(double q) const {
  if (v_.empty()) return 0.0;
  std::vector<double> s = v_;
  std::sort(s.begin(), s.end());
  const double r = std::floor(q * (double)(s.size() - 1) + 0.5);
  return s[(size_t)std::min<double>((double)s.size() - 1, r)];
}

double LatencyBook::max() const { return v_.empty() ? 0.0 : *std::max_element(v_.begin(), v_.end()); }

// The device is modelled as finishing samples at a fixed rate in admission
// order (the server hands samples out from one global range in admission
// order); a request completes when every sample up to the end of its range
// is finished, and its block returns to the arena at that moment.  Events
// (arrivals, completions) are taken in time order, an arrival before a
// completion at the same time; after each, the queue is admitted in order.
int serve_simulate(const omr_sim_request *req, size_t n, const omr_sim_config &cfg, omr_sim_outcome *out,
                   omr_sim_summary *summary) {
  if ((n && (!req || !out)) || !summary) return OMR_ERR_INVALID;
  if (!(cfg.samples_per_ms > 0.0) || !std::isfinite(cfg.samples_per_ms)) return OMR_ERR_INVALID;
  if (cfg.max_live_jobs == 0 || cfg.max_live_jobs > kJobRing) return OMR_ERR_INVALID;
  for (size_t i = 0; i < n; i++) {
    const omr_sim_request &r = req[i];
    if (!std::isfinite(r.arrival_ms) || r.arrival_ms < 0.0 || (i && r.arrival_ms < req[i - 1].arrival_ms))
      return OMR_ERR_INVALID;
    if (r.width == 0 || r.height == 0 || r.supersamples == 0 || r.spp == 0) return OMR_ERR_INVALID;
  }
  struct S {
    size_t i;
    Placement p;
    uint64_t samples, end;
  };
  std::vector<S> all(n);
  std::deque<S *> queue, live;
  Arena arena;
  arena.reset(cfg.arena_bytes);
  LatencyBook book;
  const double R = cfg.samples_per_ms;
  double now = 0.0, processed = 0.0;
  uint64_t total = 0, used = 0;
  omr_sim_summary sm{};
  for (size_t i = 0; i < n; i++) {
    out[i] = omr_sim_outcome{};
    out[i].admit_ms = out[i].done_ms = out[i].latency_ms = -1.0;
  }
  auto admit = [&]() {
    admit_in_order(
        queue, arena, cfg.max_live_jobs - live.size(), [&](S *s) { return req[s->i].cancel != 0; },
        [&](S *s) {
          out[s->i].state = 3;
          sm.jobs_cancelled++;
        },
        [&](S *s) { return &s->p; },
        [&](S *s) {
          s->end = total + s->samples;
          total = s->end;
          live.push_back(s);
          used += s->p.bytes();
          sm.peak_used_bytes = std::max<uint64_t>(sm.peak_used_bytes, used);
          sm.peak_live_jobs = std::max<uint32_t>(sm.peak_live_jobs, (uint32_t)live.size());
          omr_sim_outcome &o = out[s->i];
          o.state = 1;
          o.admit_ms = now;
          o.off = s->p.off;
          o.bytes = s->p.bytes();
          o.records = s->p.records ? 1u : 0u;
        });
  };
  const double inf = std::numeric_limits<double>::infinity();
  size_t next = 0;
  for (;;) {
    const double t_arr = next < n ? req[next].arrival_ms : inf;
    const double t_done = live.empty() ? inf : now + ((double)live.front()->end - processed) / R;
    if (t_arr == inf && t_done == inf) break;
    if (t_arr <= t_done) {
      processed = std::min((double)total, processed + (t_arr - now) * R);
      now = t_arr;
      const omr_sim_request &r = req[next];
      S &s = all[next];
      s.i = next;
      s.samples = (uint64_t)r.width * r.height * r.supersamples * r.supersamples * r.spp;
      if (!plan_request(image_bytes(r.width, r.height), s.samples, cfg.record_limit_bytes, arena.size(), &s.p)) {
        out[next].state = 5;
        sm.jobs_rejected++;
      } else {
        queue.push_back(&s);
      }
      next++;
    } else {
      S *s = live.front();
      live.pop_front();
      now = t_done;
      processed = (double)s->end;
      arena.give(s->p.off, s->p.bytes());
      used -= s->p.bytes();
      omr_sim_outcome &o = out[s->i];
      o.state = 2;
      o.done_ms = now;
      o.latency_ms = now - req[s->i].arrival_ms;
      book.add(o.latency_ms);
      sm.jobs_done++;
    }
    admit();
  }
  // requests still queued (cancelled ones behind nothing) were handled by
  // the last admission; none can remain, since an empty window admits the
  // head whenever the arena is whole and every planned request fits it
  sm.lat_p50_ms = book.percentile(0.50);
  sm.lat_p99_ms = book.percentile(0.99);
  sm.lat_max_ms = book.max();
  sm.makespan_ms = now;
  *summary = sm;
  return queue.empty() ? OMR_OK : OMR_ERR_STATE;
}

}  // namespace omr
