// Fork-join helper for the host setup: the node classification, the line
// scans that find the crossings and solve their MIB systems, the device
// table and the stage records all split into independent chunks of lines or
// planes.  Each chunk writes its own output; the caller stitches the chunks
// back together in line order, so results do not depend on the thread count.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstdlib>
#include <exception>
#include <thread>
#include <vector>

namespace mts {

// Threads one setup may use: MTS_SETUP_THREADS if set (>= 1), else the
// hardware threads, at most 16.  A thread that runs several setups side by
// side (the server's setup threads) lowers its own share with
// set_setup_thread_share.
inline int &setup_thread_share() {
  thread_local int share = 0;  // 0: not set on this thread
  return share;
}

inline void set_setup_thread_share(int n) { setup_thread_share() = n > 0 ? n : 0; }

inline int setup_threads() {
  if (setup_thread_share() > 0) return setup_thread_share();
  if (const char *ev = std::getenv("MTS_SETUP_THREADS")) {
    const int v = std::atoi(ev);
    if (v >= 1) return std::min(v, 256);
  }
  const int hw = (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(hw, 16));
}

// First item of chunk c when n items are split into `chunks` ranges.
inline int chunk_begin(int n, int chunks, int c) { return (int)((long long)n * c / chunks); }

// Chunks for n items on `threads` threads: a few per thread, so that lines
// far from the interface (no crossings) and lines through it even out.
inline int chunk_count(int n, int threads) {
  return threads <= 1 ? 1 : std::max(1, std::min(4 * threads, n));
}

// Runs f(chunk, begin, end) over [0, n) split into `chunks` (from
// chunk_count) contiguous, nearly equal ranges, on up to `threads` threads
// (the calling thread one of them) that take chunks in turn.  The first
// exception a chunk throws, in chunk order, is rethrown after every chunk has
// ended.
template <class F>
void parallel_chunks(int n, int chunks, int threads, F &&f) {
  if (chunks <= 1 || threads <= 1) {
    for (int c = 0; c < std::max(chunks, 1); c++)
      f(c, chunk_begin(n, std::max(chunks, 1), c), chunk_begin(n, std::max(chunks, 1), c + 1));
    return;
  }
  std::vector<std::exception_ptr> errs((size_t)chunks);
  std::atomic<int> next{0};
  auto work = [&]() {
    for (int c = next.fetch_add(1); c < chunks; c = next.fetch_add(1)) {
      try {
        f(c, chunk_begin(n, chunks, c), chunk_begin(n, chunks, c + 1));
      } catch (...) {
        errs[(size_t)c] = std::current_exception();
      }
    }
  };
  std::vector<std::thread> pool;
  try {
    const int extra = std::min(threads, chunks) - 1;
    pool.reserve((size_t)extra);
    for (int t = 0; t < extra; t++) pool.emplace_back(work);
  } catch (...) {
    // fewer threads than asked: the ones started and this one take every chunk
  }
  work();
  for (auto &t : pool) t.join();
  for (auto &e : errs)
    if (e) std::rethrow_exception(e);
}

}  // namespace mts
