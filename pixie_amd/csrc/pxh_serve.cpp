// pxh_serve.cpp -- the served form behind the C ABI: continuous batching of
// scene requests over the library's execution lanes.
//
// Pixie is a library with no server; this is the serving form of its raster
// path.  A request is a recorded pxh_scene rendered onto a fresh transparent
// canvas of its own size, optionally blurred.  Requests queue in arrival
// order.  One worker thread per lane takes the oldest queued request as soon
// as its canvas fits in the resident HBM arena (one device allocation carved
// first-fit), renders it on its lane, and marks it done; requests that arrive
// while others run are admitted as lanes and memory free up, and their
// kernels overlap on the device.  A finished canvas stays in the arena until
// the caller releases it.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pixie_hip.h"

struct pxh_server {
  struct Req {
    pxh_scene* scene = nullptr;
    int32_t w = 0, h = 0;
    float blur = 0;
    int state = 0;  // 0 queued, 1 running, 2 done
    int32_t status = 0;
    std::string message;
    size_t off = 0, bytes = 0;
    double tEnqueue = 0, tAdmit = 0, tDone = 0;
    pxh_image* view = nullptr;
  };
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int64_t> queue;   // queued tickets, oldest first
  std::map<int64_t, Req> reqs; // node-based: references stay valid
  int64_t nextTicket = 0;
  bool stop = false;
  char* arena = nullptr;
  size_t arenaBytes = 0;
  std::vector<std::pair<size_t, size_t>> freeList;  // (offset, length), sorted
  std::vector<std::thread> workers;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  bool alloc(size_t n, size_t* off) {
    n = (n + 255) & ~(size_t)255;
    for (size_t i = 0; i < freeList.size(); i++) {
      if (freeList[i].second >= n) {
        *off = freeList[i].first;
        freeList[i].first += n;
        freeList[i].second -= n;
        if (freeList[i].second == 0) freeList.erase(freeList.begin() + (long)i);
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = (n + 255) & ~(size_t)255;
    freeList.push_back({off, n});
    std::sort(freeList.begin(), freeList.end());
    std::vector<std::pair<size_t, size_t>> merged;
    for (auto& f : freeList) {
      if (!merged.empty() && merged.back().first + merged.back().second == f.first)
        merged.back().second += f.second;
      else
        merged.push_back(f);
    }
    freeList.swap(merged);
  }

  void worker(int lane) {
    for (;;) {
      int64_t id;
      Req* r;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          if (stop) return;
          if (!queue.empty()) {
            Req& head = reqs[queue.front()];
            size_t off;
            if (alloc(head.bytes, &off)) {  // FIFO: only the oldest may go
              head.off = off;
              break;
            }
          }
          cv.wait(lk);
        }
        id = queue.front();
        queue.pop_front();
        r = &reqs[id];
        r->state = 1;
        r->tAdmit = now();
        cv.notify_all();  // the next queued request may fit too
      }
      pxh_image* img = pxh_image_view(arena + r->off, r->w, r->h, PXH_FORMAT_RGBA8);
      int32_t st = img ? 0 : PXH_ERR_ARG;
      std::string msg;
      if (!st) st = pxh_image_fill_on(img, 0u, lane);
      if (!st) st = pxh_scene_render_on(r->scene, img, lane);
      if (!st && r->blur != 0.0f) st = pxh_blur_on(img, r->blur, 0u, lane);
      if (st) msg = pxh_last_error();
      // drain the lane whatever happened, so that nothing of this request
      // is still running when its canvas is handed back
      const int32_t st2 = pxh_lane_synchronize(lane);
      if (!st && st2) {
        st = st2;
        msg = pxh_last_error();
      }
      std::lock_guard<std::mutex> lk(mu);
      r->status = st;
      r->message = msg;
      r->view = img;
      r->state = 2;
      r->tDone = now();
      cv.notify_all();
    }
  }
};

extern "C" {

pxh_server* pxh_server_new(int32_t n_lanes, int64_t hbm_budget) {
  if (pxh_init() != 0) return nullptr;
  if (n_lanes < 1 || n_lanes > pxh_lanes() - 1 || hbm_budget < 256) return nullptr;
  pxh_server* s = new pxh_server();
  s->arenaBytes = ((size_t)hbm_budget + 255) & ~(size_t)255;
  if (hipMalloc((void**)&s->arena, s->arenaBytes) != hipSuccess) {
    delete s;
    return nullptr;
  }
  s->freeList.push_back({0, s->arenaBytes});
  for (int k = 0; k < n_lanes; k++) s->workers.emplace_back([s, k] { s->worker(k + 1); });
  return s;
}

void pxh_server_free(pxh_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  for (auto& t : s->workers) t.join();  // running requests finish first
  for (auto& kv : s->reqs)
    if (kv.second.view) pxh_image_free(kv.second.view);
  (void)hipFree(s->arena);
  delete s;
}

int64_t pxh_server_enqueue(pxh_server* s, pxh_scene* scene, int32_t width, int32_t height,
                           float blur_radius) {
  if (!s || !scene || width <= 0 || height <= 0) return -PXH_ERR_ARG;
  const size_t bytes = (size_t)width * (size_t)height * 4;
  if (bytes > s->arenaBytes) return -PXH_ERR_CAPACITY;
  std::lock_guard<std::mutex> lk(s->mu);
  const int64_t id = s->nextTicket++;
  pxh_server::Req& r = s->reqs[id];
  r.scene = scene;
  r.w = width;
  r.h = height;
  r.blur = blur_radius;
  r.bytes = bytes;
  r.tEnqueue = s->now();
  s->queue.push_back(id);
  s->cv.notify_all();
  return id;
}

int pxh_server_poll(pxh_server* s, int64_t ticket, int32_t* done, int32_t* status) {
  if (!s) return PXH_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return PXH_ERR_ARG;
  if (done) *done = it->second.state == 2 ? 1 : 0;
  if (status) *status = it->second.state == 2 ? it->second.status : 0;
  return 0;
}

int pxh_server_wait(pxh_server* s, int64_t ticket, int32_t* status) {
  if (!s) return PXH_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return PXH_ERR_ARG;
  s->cv.wait(lk, [&] { return it->second.state == 2; });
  if (status) *status = it->second.status;
  return 0;
}

pxh_image* pxh_server_image(pxh_server* s, int64_t ticket) {
  if (!s) return nullptr;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end() || it->second.state != 2) return nullptr;
  return it->second.view;
}

const char* pxh_server_message(pxh_server* s, int64_t ticket) {
  if (!s) return "";
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return "";
  return it->second.message.c_str();
}

int pxh_server_times(pxh_server* s, int64_t ticket, double* enqueued, double* admitted,
                     double* done) {
  if (!s) return PXH_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end()) return PXH_ERR_ARG;
  if (enqueued) *enqueued = it->second.tEnqueue;
  if (admitted) *admitted = it->second.tAdmit;
  if (done) *done = it->second.tDone;
  return 0;
}

double pxh_server_now(pxh_server* s) { return s ? s->now() : 0.0; }

int pxh_server_release(pxh_server* s, int64_t ticket) {
  if (!s) return PXH_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(ticket);
  if (it == s->reqs.end() || it->second.state != 2) return PXH_ERR_ARG;
  if (it->second.view) pxh_image_free(it->second.view);
  s->release(it->second.off, it->second.bytes);
  s->reqs.erase(it);
  s->cv.notify_all();
  return 0;
}

}  // extern "C"
