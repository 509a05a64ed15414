// serve.hip -- a checkpoint-serving loop over the batched pipeline.
//
// Clients register once and then submit checkpoints of their own buffer; each
// client's Merkle tree and digest table stay resident in HBM between its
// checkpoints.  Submissions wait in one arrival-ordered queue.  Each lane (a
// worker thread with its own stream) repeatedly takes every queued request it may
// run -- at most one per client, in arrival order, and for a client without
// resident state only if its estimated state fits the memory budget -- and runs
// them as one batched step.  Requests that arrive while lanes are busy join the
// next step of whichever lane frees first, and steps of different lanes overlap on
// the device.  A client's checkpoints run in submission order, so every result
// equals the client's serial checkpoint sequence.  After posting a step's results
// a lane grows any of its clients' digest tables that the next checkpoint could
// not run in without pausing, so that rehash falls outside requests; grown tables
// count against the budget.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "internal.hpp"

namespace mtc {
void checkpoint_batch_impl(mtc_dedup* const* hs, const uint8_t* const* datas, const uint64_t* lens,
                           const int* baselines, uint32_t count, StepCtx& ctx, hipStream_t st, bool use_graph);
void pregrow(mtc_dedup* h, hipStream_t st);
}  // namespace mtc

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  uint32_t client = 0;
  const uint8_t* data = nullptr;
  uint64_t len = 0;
  int baseline = 0;
  hipEvent_t ready = nullptr;  // inputs written (recorded on the submitter's stream)
  double t_submit = 0, t_start = 0, t_done = 0;
  bool done = false;
  int status = MTC_OK;
  std::string error;
  mtc_header header{};
  uint64_t diff_size = 0;
};

struct Client {
  mtc_dedup* h = nullptr;
  uint64_t max_len = 0;      // longest buffer the client may submit (its state estimate's size)
  uint64_t state_est = 0;
  uint64_t table_slots = 0;  // table slots counted in state_est
  uint64_t table_extra = 0;  // bytes of table grown beyond them (counted while resident)
  bool resident = false;
  bool busy = false;
  bool removed = false;
};

struct Lane {
  hipStream_t stream = nullptr;
  mtc::StepCtx step;
  std::thread thread;
};

}  // namespace

struct mtc_server {
  int device = 0;
  uint64_t budget = 0;
  uint32_t max_batch = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::vector<Client> clients;
  std::deque<uint64_t> pending;  // tickets in arrival order
  std::unordered_map<uint64_t, Request> reqs;
  uint64_t next_ticket = 1;
  uint64_t resident_bytes = 0;
  bool stop = false;
  std::vector<Lane*> lanes;
  std::vector<hipEvent_t> free_events;
  // statistics
  uint64_t steps = 0, served = 0, bytes = 0;
  std::string last_error;

  // Picks the next step's requests (lock held).
  std::vector<uint64_t> pick() {
    std::vector<uint64_t> out;
    std::unordered_set<uint32_t> blocked;
    for (auto it = pending.begin(); it != pending.end() && out.size() < max_batch;) {
      Request& r = reqs[*it];
      Client& c = clients[r.client];
      bool take = !c.busy && !blocked.count(r.client);
      if (take && !c.resident) {
        if (resident_bytes + c.state_est <= budget) {
          c.resident = true;
          resident_bytes += c.state_est;
        } else {
          take = false;
        }
      }
      if (!take) {
        blocked.insert(r.client);  // later requests of this client wait behind this one
        ++it;
        continue;
      }
      c.busy = true;
      out.push_back(*it);
      it = pending.erase(it);
    }
    return out;
  }

  void run_lane(Lane* lane) {
    (void)hipSetDevice(device);
    std::vector<uint64_t> tickets;
    std::vector<mtc_dedup*> hs;
    std::vector<const uint8_t*> ds;
    std::vector<uint64_t> ls;
    std::vector<int> bs;
    while (true) {
      {
        std::unique_lock<std::mutex> lk(mu);
        tickets.clear();
        cv_work.wait(lk, [&] {
          if (stop) return true;
          tickets = pick();
          return !tickets.empty();
        });
        if (tickets.empty()) return;  // stopping
        hs.clear();
        ds.clear();
        ls.clear();
        bs.clear();
        const double t = now_s();
        for (uint64_t tk : tickets) {
          Request& r = reqs[tk];
          r.t_start = t;
          hs.push_back(clients[r.client].h);
          ds.push_back(r.data);
          ls.push_back(r.len);
          bs.push_back(r.baseline);
          if (r.ready) (void)hipStreamWaitEvent(lane->stream, r.ready, 0);
        }
      }
      int status = MTC_OK;
      std::string err;
      try {
        mtc::checkpoint_batch_impl(hs.data(), ds.data(), ls.data(), bs.data(), (uint32_t)hs.size(), lane->step,
                                   lane->stream, false);
      } catch (const mtc::Status& s) {
        status = s.code;
        err = s.msg;
      } catch (const mtc::HipError& e) {
        status = (e.err == hipErrorOutOfMemory) ? MTC_ERR_OUT_OF_MEMORY : MTC_ERR_HIP;
        err = std::string(e.what) + ": " + hipGetErrorString(e.err);
      } catch (...) {
        status = MTC_ERR_HIP;
        err = "unexpected exception";
      }
      std::vector<uint32_t> served_clients;
      {
        std::lock_guard<std::mutex> lk(mu);
        const double t = now_s();
        steps++;
        for (size_t i = 0; i < tickets.size(); i++) {
          Request& r = reqs[tickets[i]];
          Client& c = clients[r.client];
          served_clients.push_back(r.client);
          r.t_done = t;
          r.done = true;
          r.status = status;
          r.error = err;
          if (status == MTC_OK) {
            r.header = c.h->last_header;
            r.diff_size = c.h->last_diff_size;
            c.h->current_id += 1;
            served++;
            bytes += r.len;
          }
          if (r.ready) free_events.push_back(r.ready);
          r.ready = nullptr;
        }
      }
      cv_done.notify_all();
      // Results are posted; the clients stay busy while a table that their next
      // checkpoint would have to grow is grown here, outside their requests.  A
      // failure leaves the growth to that checkpoint.
      if (status == MTC_OK)
        for (mtc_dedup* h : hs) {
          try {
            mtc::pregrow(h, lane->stream);
          } catch (...) {
          }
        }
      {
        // (a posted request may already have been collected and erased)
        std::lock_guard<std::mutex> lk(mu);
        for (uint32_t ci : served_clients) {
          Client& c = clients[ci];
          // grown tables count against the budget
          const uint64_t slots = c.h->table.count;
          const uint64_t extra = slots > c.table_slots ? 32 * (slots - c.table_slots) : 0;
          if (c.resident) resident_bytes = resident_bytes - c.table_extra + extra;
          c.table_extra = extra;
          c.busy = false;
        }
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
};

static thread_local std::string g_serve_error;

template <class F>
static int serve_guarded(mtc_server* s, F&& body) {
  int code = MTC_OK;
  std::string msg;
  try {
    body();
  } catch (const mtc::Status& e) {
    code = e.code;
    msg = e.msg;
  } catch (const mtc::HipError& e) {
    code = (e.err == hipErrorOutOfMemory) ? MTC_ERR_OUT_OF_MEMORY : MTC_ERR_HIP;
    msg = std::string(e.what) + ": " + hipGetErrorString(e.err);
  } catch (const std::bad_alloc&) {
    code = MTC_ERR_OUT_OF_MEMORY;
    msg = "host allocation failed";
  }
  if (code != MTC_OK) {
    g_serve_error = msg;
    if (s) {
      std::lock_guard<std::mutex> lk(s->mu);
      s->last_error = msg;
    }
  }
  return code;
}

extern "C" {

int mtc_server_create(mtc_server** out, int device, uint64_t memory_budget, uint32_t lanes, uint32_t max_batch) {
  if (!out || lanes == 0 || lanes > 8 || max_batch == 0 || max_batch > mtc::kMaxBatch || memory_budget == 0)
    return MTC_ERR_INVALID_ARG;
  *out = nullptr;
  mtc_server* s = new mtc_server();
  s->device = device;
  s->budget = memory_budget;
  s->max_batch = max_batch;
  const int rc = serve_guarded(nullptr, [&] {
    MTC_HIP(hipSetDevice(device));
    for (uint32_t i = 0; i < lanes; i++) {
      Lane* l = new Lane();
      s->lanes.push_back(l);
      MTC_HIP(hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking));
    }
    for (Lane* l : s->lanes) l->thread = std::thread([s, l] { s->run_lane(l); });
  });
  if (rc != MTC_OK) {
    mtc_server_destroy(s);
    return rc;
  }
  *out = s;
  return MTC_OK;
}

int mtc_server_destroy(mtc_server* s) {
  if (!s) return MTC_OK;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // Finish what was admitted; requests still queued are dropped.
    s->cv_done.wait(lk, [&] {
      for (const auto& c : s->clients)
        if (c.busy) return false;
      return true;
    });
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (Lane* l : s->lanes)
    if (l->thread.joinable()) l->thread.join();
  (void)hipSetDevice(s->device);
  for (Lane* l : s->lanes) {
    if (l->stream) (void)hipStreamSynchronize(l->stream);
    l->step.reset_graph();
    if (l->stream) (void)hipStreamDestroy(l->stream);
    delete l;
  }
  for (auto& kv : s->reqs)
    if (kv.second.ready) (void)hipEventDestroy(kv.second.ready);
  for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
  for (auto& c : s->clients)
    if (c.h) mtc_destroy(c.h);
  delete s;
  return MTC_OK;
}

int mtc_server_add_client(mtc_server* s, uint32_t chunk_size, int hash_kind, uint64_t max_len, uint32_t* client) {
  if (!s || !client || max_len == 0 || chunk_size == 0) return MTC_ERR_INVALID_ARG;
  // A client whose resident state could never fit the budget would wait forever.
  const uint64_t n = (max_len + chunk_size - 1) / chunk_size;
  const double lf = mtc::Tune{}.load_factor;
  if (n > (1ull << 31)) return MTC_ERR_INVALID_ARG;
  if (mtc::table_slots(n, lf) > mtc::kMaxTableSlots) return MTC_ERR_TABLE_FULL;
  const uint64_t est = mtc::state_bytes(n, chunk_size, lf);
  if (est > s->budget) {
    g_serve_error = "the client's resident state (" + std::to_string(est) + " bytes) exceeds the memory budget";
    std::lock_guard<std::mutex> lk(s->mu);
    s->last_error = g_serve_error;
    return MTC_ERR_OUT_OF_MEMORY;
  }
  mtc_dedup* h = nullptr;
  int rc = mtc_create(&h, chunk_size, MTC_APPROACH_TREE, hash_kind, s->device);
  if (rc != MTC_OK) return rc;
  std::lock_guard<std::mutex> lk(s->mu);
  Client c;
  c.h = h;
  c.max_len = max_len;
  c.state_est = est;
  c.table_slots = mtc::table_slots(n, lf) + 1;
  s->clients.push_back(c);
  *client = (uint32_t)(s->clients.size() - 1);
  return MTC_OK;
}

int mtc_server_client(mtc_server* s, uint32_t client, mtc_dedup** handle) {
  if (!s || !handle) return MTC_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (client >= s->clients.size() || s->clients[client].removed) return MTC_ERR_INVALID_ARG;
  *handle = s->clients[client].h;
  return MTC_OK;
}

int mtc_server_remove_client(mtc_server* s, uint32_t client) {
  if (!s) return MTC_ERR_INVALID_ARG;
  mtc_dedup* h = nullptr;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    if (client >= s->clients.size() || s->clients[client].removed) return MTC_ERR_INVALID_ARG;
    s->cv_done.wait(lk, [&] {
      if (s->clients[client].busy) return false;
      for (uint64_t t : s->pending)
        if (s->reqs[t].client == client) return false;
      return true;
    });
    Client& c = s->clients[client];
    if (c.resident) s->resident_bytes -= c.state_est + c.table_extra;
    c.resident = false;
    c.removed = true;
    h = c.h;
    c.h = nullptr;
  }
  mtc_destroy(h);
  s->cv_work.notify_all();  // freed memory may admit waiting clients
  return MTC_OK;
}

int mtc_server_submit(mtc_server* s, uint32_t client, const void* dev_data, uint64_t len, int make_baseline,
                      void* stream, uint64_t* ticket) {
  if (!s || !dev_data || len == 0 || !ticket) return MTC_ERR_INVALID_ARG;
  return serve_guarded(s, [&] {
    MTC_HIP(hipSetDevice(s->device));
    hipEvent_t ev = nullptr;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (client >= s->clients.size() || s->clients[client].removed)
        throw mtc::Status{MTC_ERR_INVALID_ARG, "unknown client"};
      if (len > s->clients[client].max_len)
        throw mtc::Status{MTC_ERR_INVALID_ARG, "submission longer than the client's registered max_len"};
      if (!s->free_events.empty()) {
        ev = s->free_events.back();
        s->free_events.pop_back();
      }
    }
    if (!ev) MTC_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    MTC_HIP(hipEventRecord(ev, static_cast<hipStream_t>(stream)));
    {
      std::lock_guard<std::mutex> lk(s->mu);
      const uint64_t t = s->next_ticket++;
      Request& r = s->reqs[t];
      r.client = client;
      r.data = static_cast<const uint8_t*>(dev_data);
      r.len = len;
      r.baseline = make_baseline ? 1 : 0;
      r.ready = ev;
      r.t_submit = now_s();
      s->pending.push_back(t);
      *ticket = t;
    }
    s->cv_work.notify_one();
  });
}

// Result of a request: MTC_PENDING while it has not finished (wait == 0), else its
// status; a finished request's record is released by this call.  Times are
// CLOCK_MONOTONIC seconds.
int mtc_server_result(mtc_server* s, uint64_t ticket, int wait, mtc_header* header, uint64_t* diff_size,
                      double* t_submit, double* t_start, double* t_done) {
  if (!s) return MTC_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return MTC_ERR_INVALID_ARG;
  if (!it->second.done) {
    if (!wait) return MTC_PENDING;
    s->cv_done.wait(lk, [&] { return s->reqs[ticket].done; });
    it = s->reqs.find(ticket);
  }
  const Request& r = it->second;
  if (header) *header = r.header;
  if (diff_size) *diff_size = r.diff_size;
  if (t_submit) *t_submit = r.t_submit;
  if (t_start) *t_start = r.t_start;
  if (t_done) *t_done = r.t_done;
  const int st = r.status;
  if (st != MTC_OK) s->last_error = r.error;
  s->reqs.erase(it);
  return st;
}

int mtc_server_stats(mtc_server* s, uint64_t* steps, uint64_t* served, uint64_t* bytes, uint64_t* resident_bytes,
                     uint32_t* queued) {
  if (!s) return MTC_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  if (steps) *steps = s->steps;
  if (served) *served = s->served;
  if (bytes) *bytes = s->bytes;
  if (resident_bytes) *resident_bytes = s->resident_bytes;
  if (queued) *queued = (uint32_t)s->pending.size();
  return MTC_OK;
}

const char* mtc_server_last_error(mtc_server* s) {
  if (!s) return g_serve_error.c_str();
  std::lock_guard<std::mutex> lk(s->mu);
  thread_local std::string copy;
  copy = s->last_error;
  return copy.c_str();
}

}  // extern "C"
