// serve.hip -- continuous-batching solve server over a swalbe_batch (host
// code only).  Requests are admitted in submission order while their film
// fits the arena and a slot is free; each poll enqueues one tick of steps
// for every resident film and hands back the films that finished in an
// earlier tick, so downloads overlap the current tick on the device.
#include <algorithm>
#include <deque>
#include <map>
#include <vector>

#include "swalbe_internal.h"

using namespace swalbe;

namespace {

struct Pending {
  int64_t id;
  swalbe_params p;
  int32_t Lx, Ly;
  swalbe_film_init init;
  int64_t nsteps;
};

struct Resident {
  int64_t id;
  int32_t Lx, Ly;
  int64_t remaining;        // steps not yet enqueued
  int64_t finished_at = -1; // steps_enqueued when the last step was enqueued
};

struct Done {
  int64_t id;
  int32_t Lx, Ly;
  std::vector<double> h;
};

int bad(const char* msg) {
  set_error(msg);
  return SWALBE_EINVAL;
}

}  // namespace

struct swalbe_server {
  swalbe_batch* batch = nullptr;
  int32_t max_films = 0;
  int64_t tick = 0;
  int64_t next_id = 0;
  int64_t steps_enqueued = 0;
  std::deque<Pending> queue;
  std::map<int32_t, Resident> resident;  // slot -> film
  std::deque<Done> done;
};

namespace {

int admit(swalbe_server* s) {
  while (!s->queue.empty() && (int32_t)s->resident.size() < s->max_films) {
    const Pending& q = s->queue.front();
    size_t need = 0, freeb = 0;
    if (int rc = swalbe_batch_film_bytes(q.Lx, q.Ly, q.p.tau, &need)) return rc;
    if (int rc = swalbe_batch_stats(s->batch, nullptr, nullptr, nullptr, &freeb)) return rc;
    if (need > freeb) break;
    int32_t slot = -1;
    const int rc = swalbe_batch_add(s->batch, &q.p, q.Lx, q.Ly, &q.init, q.nsteps, &slot);
    if (rc == SWALBE_ENOMEM) break;  // free bytes fragmented: wait for a retirement
    if (rc) return rc;
    Resident r{q.id, q.Lx, q.Ly, q.nsteps};
    if (q.nsteps == 0) r.finished_at = s->steps_enqueued;
    s->resident[slot] = r;
    s->queue.pop_front();
  }
  return SWALBE_OK;
}

int retire(swalbe_server* s, int64_t upto) {
  for (auto it = s->resident.begin(); it != s->resident.end();) {
    const Resident& r = it->second;
    if (r.finished_at < 0 || r.finished_at > upto) {
      ++it;
      continue;
    }
    Done d{r.id, r.Lx, r.Ly, std::vector<double>((size_t)r.Lx * r.Ly)};
    if (int rc = swalbe_batch_download(s->batch, it->first, SWALBE_HEIGHT, d.h.data(),
                                       d.h.size() * sizeof(double)))
      return rc;
    if (int rc = swalbe_batch_remove(s->batch, it->first)) return rc;
    s->done.push_back(std::move(d));
    it = s->resident.erase(it);
  }
  return SWALBE_OK;
}

}  // namespace

extern "C" {

int swalbe_serve_create(size_t capacity_bytes, int32_t max_films, int32_t tick_steps,
                        int32_t graph_steps, swalbe_server** out) {
  if (!out || tick_steps < 1 || graph_steps < 0) return bad("swalbe_serve_create: bad arguments");
  swalbe_server* s = new swalbe_server();
  if (int rc = swalbe_batch_create(SWALBE_F32, capacity_bytes, max_films, &s->batch)) {
    delete s;
    return rc;
  }
  // requests return heights only: distributions and pressure of a film are
  // written on its last step only
  swalbe_batch_set_final_outputs_only(s->batch, 1);
  if (graph_steps > 0) swalbe_batch_set_graph(s->batch, graph_steps);
  s->max_films = max_films;
  s->tick = tick_steps;
  *out = s;
  return SWALBE_OK;
}

int swalbe_serve_destroy(swalbe_server* s) {
  if (!s) return SWALBE_OK;
  swalbe_batch_destroy(s->batch);
  delete s;
  return SWALBE_OK;
}

int swalbe_serve_submit(swalbe_server* s, const swalbe_params* p, int32_t Lx, int32_t Ly,
                        const swalbe_film_init* init, int64_t nsteps, int64_t* request_id) {
  if (!s || !p || !init || !request_id || Lx < 1 || Ly < 1 || nsteps < 0)
    return bad("swalbe_serve_submit: bad arguments");
  size_t need = 0;
  if (int rc = swalbe_batch_film_bytes(Lx, Ly, p->tau, &need)) return rc;
  size_t used = 0, freeb = 0;
  swalbe_batch_stats(s->batch, nullptr, nullptr, &used, &freeb);
  if (need > used + freeb) {
    set_error("swalbe_serve_submit: the film is larger than the whole arena");
    return SWALBE_ENOMEM;
  }
  s->queue.push_back(Pending{s->next_id, *p, Lx, Ly, *init, nsteps});
  *request_id = s->next_id++;
  return SWALBE_OK;
}

int swalbe_serve_poll(swalbe_server* s, int32_t* n_done, int32_t* n_resident, int32_t* n_queued) {
  if (!s) return bad("swalbe_serve_poll: null server");
  if (int rc = admit(s)) return rc;
  int64_t k = s->tick;
  bool live = false;
  for (auto& kv : s->resident)
    if (kv.second.remaining > 0) {
      live = true;
      k = std::min(k, kv.second.remaining);
    }
  if (live) {
    const int64_t prev = s->steps_enqueued;
    if (int rc = swalbe_batch_step(s->batch, k)) return rc;
    s->steps_enqueued += k;
    for (auto& kv : s->resident) {
      Resident& r = kv.second;
      if (r.remaining <= 0) continue;
      r.remaining -= k;
      if (r.remaining <= 0) {
        r.remaining = 0;
        r.finished_at = s->steps_enqueued;
      }
    }
    // the previous tick has completed once at most this one is pending
    if (int rc = swalbe_batch_wait(s->batch, k)) return rc;
    if (int rc = retire(s, prev)) return rc;
  } else if (!s->resident.empty()) {
    if (int rc = swalbe_batch_wait(s->batch, 0)) return rc;
    if (int rc = retire(s, s->steps_enqueued)) return rc;
  }
  if (n_done) *n_done = (int32_t)s->done.size();
  if (n_resident) *n_resident = (int32_t)s->resident.size();
  if (n_queued) *n_queued = (int32_t)s->queue.size();
  return SWALBE_OK;
}

int swalbe_serve_peek(swalbe_server* s, int64_t* request_id, int32_t* Lx, int32_t* Ly) {
  if (!s || !request_id || !Lx || !Ly) return bad("swalbe_serve_peek: bad arguments");
  if (s->done.empty()) return bad("swalbe_serve_peek: no finished request");
  *request_id = s->done.front().id;
  *Lx = s->done.front().Lx;
  *Ly = s->done.front().Ly;
  return SWALBE_OK;
}

int swalbe_serve_take(swalbe_server* s, double* height, size_t bytes) {
  if (!s || !height) return bad("swalbe_serve_take: bad arguments");
  if (s->done.empty()) return bad("swalbe_serve_take: no finished request");
  Done& d = s->done.front();
  if (bytes != d.h.size() * sizeof(double))
    return bad("swalbe_serve_take: byte count does not match Lx*Ly doubles");
  std::copy(d.h.begin(), d.h.end(), height);
  s->done.pop_front();
  return SWALBE_OK;
}

}  // extern "C"
