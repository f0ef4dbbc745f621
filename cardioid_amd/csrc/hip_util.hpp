// hip_util.hpp -- error handling and device buffers for the host side.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace cdh {

struct Error : std::runtime_error {
   int code;
   Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

enum { kOk = 0, kErrArg = 1, kErrHip = 2, kErrUnsupported = 3, kErrState = 4 };

#define CDH_HIP(call)                                                                   \
   do {                                                                                 \
      hipError_t e_ = (call);                                                           \
      if (e_ != hipSuccess)                                                             \
         throw ::cdh::Error(::cdh::kErrHip, std::string(#call) + ": " + hipGetErrorString(e_) + \
                                                " (" __FILE__ ")");                     \
   } while (0)

#define CDH_REQUIRE(cond, msg)                                                          \
   do {                                                                                 \
      if (!(cond)) throw ::cdh::Error(::cdh::kErrArg, msg);                             \
   } while (0)

#define CDH_LAUNCH_CHECK() CDH_HIP(hipGetLastError())

// Owning device allocation.
template <class T>
struct DevBuf {
   T *p = nullptr;
   size_t n = 0;
   DevBuf() = default;
   explicit DevBuf(size_t count) { alloc(count); }
   DevBuf(const DevBuf &) = delete;
   DevBuf &operator=(const DevBuf &) = delete;
   ~DevBuf() { release(); }
   void alloc(size_t count)
   {
      release();
      n = count;
      if (count) CDH_HIP(hipMalloc((void **)&p, count * sizeof(T)));
   }
   void release()
   {
      if (p) (void)hipFree(p);
      p = nullptr;
      n = 0;
   }
   void upload(const T *host, size_t count, size_t offset = 0)
   {
      if (count) CDH_HIP(hipMemcpy(p + offset, host, count * sizeof(T), hipMemcpyHostToDevice));
   }
   void download(T *host, size_t count, size_t offset = 0) const
   {
      if (count) CDH_HIP(hipMemcpy(host, p + offset, count * sizeof(T), hipMemcpyDeviceToHost));
   }
   void zero()
   {
      if (n) CDH_HIP(hipMemset(p, 0, n * sizeof(T)));
   }
};

inline int grid_for(int64_t n, int block) { return (int)((n + block - 1) / block); }

void set_last_error(const std::string &msg);

// Runs f and converts exceptions into the C-ABI's error codes.
template <class F>
int guard(F &&f)
{
   try {
      f();
      return kOk;
   } catch (const Error &e) {
      set_last_error(e.what());
      return e.code;
   } catch (const std::exception &e) {
      set_last_error(e.what());
      return kErrState;
   }
}

}  // namespace cdh
