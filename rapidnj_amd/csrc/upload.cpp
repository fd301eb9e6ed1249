// Copies between the caller's pageable host memory and the device through pinned staging chunks.
//
// A copy from or to pageable memory goes through the driver's own staging at a fraction of the
// link rate (and, into freshly allocated memory, at the pace of one thread's page faults), and
// registering the caller's memory is not an option. Instead every transfer is cut into chunks of
// kChunkBytes that pass through kChunks pinned buffers, with a few library threads copying on the
// host side while the DMA of a neighbouring chunk runs:
//   upload:   threads fill chunk k, its H2D DMA is enqueued, chunk k+1 fills meanwhile; a chunk is
//             refilled only after the event behind its previous DMA has completed. upload()
//             returns as soon as every source byte has been read, so the caller may reuse its
//             buffer at once; work enqueued later on the same stream sees the data.
//   download: the D2H DMAs of the first kChunks chunks are enqueued at once; as each completes,
//             threads copy it out and the chunk takes the DMA of the next chunk not yet enqueued.
//             download() returns when every byte is in the caller's buffer.
#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <thread>

#include "runtime.hpp"

namespace rnj {

// Fixed worker threads that run the parts of one copy; the calling thread takes parts too.
class CopyPool {
 public:
  explicit CopyPool(unsigned workers) {
    for (unsigned t = 0; t < workers; ++t) threads_.emplace_back([this] { work(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : threads_) t.join();
  }
  // fn(p) for p in [0, parts); returns when every part has finished.
  void run(unsigned parts, const std::function<void(unsigned)>& fn) {
    std::unique_lock<std::mutex> lk(m_);
    fn_ = &fn;
    parts_ = parts;
    next_ = 0;
    finished_ = 0;
    ++gen_;
    cv_.notify_all();
    take(lk);
    done_.wait(lk, [this] { return finished_ == parts_; });
    fn_ = nullptr;
  }

 private:
  void take(std::unique_lock<std::mutex>& lk) {
    while (fn_ && next_ < parts_) {
      const unsigned p = next_++;
      const std::function<void(unsigned)>* fn = fn_;
      lk.unlock();
      (*fn)(p);
      lk.lock();
      if (++finished_ == parts_) done_.notify_all();
    }
  }
  void work() {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      take(lk);
    }
  }
  std::vector<std::thread> threads_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(unsigned)>* fn_ = nullptr;
  unsigned parts_ = 0, next_ = 0, finished_ = 0;
  uint64_t gen_ = 0;
  bool stop_ = false;
};

namespace {

// Bytes [off, off + m) of the row-packed stream (rows of `len` bytes) from rows at `pitch`.
void copy_span(uint8_t* dst, const char* src, size_t pitch, size_t len, size_t off, size_t m) {
  if (pitch == len) {
    std::memcpy(dst, src + off, m);
    return;
  }
  while (m) {
    const size_t r = off / len, c = off % len, take = std::min(len - c, m);
    std::memcpy(dst, src + r * pitch + c, take);
    dst += take;
    off += take;
    m -= take;
  }
}

unsigned copy_parts() {
  const unsigned hw = std::thread::hardware_concurrency();
  return std::max(1u, std::min(8u, hw / 2));
}

}  // namespace

Stager::Stager() = default;

Stager::~Stager() {
  for (int k = 0; k < kChunks; ++k) {
    if (ev_[k]) {
      (void)hipEventSynchronize(ev_[k]);
      (void)hipEventDestroy(ev_[k]);
    }
    if (chunk_[k]) (void)hipHostFree(chunk_[k]);
  }
}

hipError_t Stager::ensure() {
  hipError_t e = hipSuccess;
  for (int k = 0; k < kChunks && e == hipSuccess; ++k) {
    if (!chunk_[k]) e = hipHostMalloc((void**)&chunk_[k], kChunkBytes, hipHostMallocDefault);
    if (e == hipSuccess && !ev_[k]) e = hipEventCreateWithFlags(&ev_[k], hipEventDisableTiming);
  }
  if (e == hipSuccess && !pool_) {
    parts_ = copy_parts();
    pool_ = std::make_unique<CopyPool>(parts_ - 1);
  }
  return e;
}

hipError_t Stager::upload(uint8_t* dst, const char* src, size_t pitch, size_t len, size_t rows,
                            hipStream_t stream) {
  const size_t total = len * rows;
  if (total == 0) return hipSuccess;
  std::lock_guard<std::mutex> lock(mu_);
  hipError_t e = ensure();
  for (size_t off = 0; e == hipSuccess && off < total; off += kChunkBytes) {
    const size_t m = std::min(kChunkBytes, total - off);
    const int k = next_;
    next_ = (next_ + 1) % kChunks;
    e = hipEventSynchronize(ev_[k]);  // the chunk's previous DMA has read it
    if (e != hipSuccess) break;
    uint8_t* buf = chunk_[k];
    const unsigned parts = m >= kParallelBytes ? parts_ : 1;
    const std::function<void(unsigned)> part = [&](unsigned p) {
      const size_t a = m * p / parts, b = m * (p + 1) / parts;
      copy_span(buf + a, src, pitch, len, off + a, b - a);
    };
    if (parts == 1) part(0);
    else pool_->run(parts, part);
    e = hipMemcpyAsync(dst + off, buf, m, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipEventRecord(ev_[k], stream);
  }
  return e;
}

hipError_t Stager::download(void* dst_host, const void* src_dev, size_t bytes, hipStream_t stream) {
  if (bytes == 0) return hipSuccess;
  std::lock_guard<std::mutex> lock(mu_);
  hipError_t e = ensure();
  uint8_t* dst = static_cast<uint8_t*>(dst_host);
  const uint8_t* src = static_cast<const uint8_t*>(src_dev);
  const size_t nchunks = (bytes + kChunkBytes - 1) / kChunkBytes;
  auto enqueue = [&](size_t c) {  // chunk c into buffer c % kChunks
    const size_t off = c * kChunkBytes, m = std::min(kChunkBytes, bytes - off);
    const int k = (int)(c % kChunks);
    hipError_t r = hipEventSynchronize(ev_[k]);  // an earlier upload's DMA may still read it
    if (r == hipSuccess) r = hipMemcpyAsync(chunk_[k], src + off, m, hipMemcpyDeviceToHost, stream);
    if (r == hipSuccess) r = hipEventRecord(ev_[k], stream);
    return r;
  };
  size_t issued = 0;
  for (; e == hipSuccess && issued < std::min(nchunks, (size_t)kChunks); ++issued) e = enqueue(issued);
  for (size_t c = 0; e == hipSuccess && c < nchunks; ++c) {
    const int k = (int)(c % kChunks);
    e = hipEventSynchronize(ev_[k]);
    if (e != hipSuccess) break;
    const size_t off = c * kChunkBytes, m = std::min(kChunkBytes, bytes - off);
    const uint8_t* buf = chunk_[k];
    const unsigned parts = m >= kParallelBytes ? parts_ : 1;
    const std::function<void(unsigned)> part = [&](unsigned p) {
      const size_t a = m * p / parts, b = m * (p + 1) / parts;
      std::memcpy(dst + off + a, buf + a, b - a);
    };
    if (parts == 1) part(0);
    else pool_->run(parts, part);
    if (issued < nchunks) e = enqueue(issued++);
  }
  if (e != hipSuccess) (void)hipStreamSynchronize(stream);  // no DMA left writing the chunks
  return e;
}

}  // namespace rnj
