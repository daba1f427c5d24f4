// cudaStatus.h -- CudaStatus / CudaStatusOr<T> over hipError_t.
//
// Same names, constructors and accessors as hm-cupano's src/cuda/cudaStatus.h
// (CudaStatus :12-75, CudaStatusOr :80-193, CUDA_RETURN_IF_ERROR :203-209);
// the error type is HIP's.
#pragma once

#include <hip/hip_runtime_api.h>

#include <optional>
#include <ostream>
#include <stdexcept>
#include <string>
#include <utility>

class CudaStatus {
 public:
  CudaStatus() = default;
  CudaStatus(hipError_t code) : code_(code), message_(code == hipSuccess ? "OK" : hipGetErrorString(code)) {}
  CudaStatus(hipError_t code, std::string message)
      : code_(code),
        message_(code == hipSuccess ? std::move(message) : std::string(hipGetErrorString(code)) + ": " + message) {}

  bool ok() const { return code_ == hipSuccess; }
  hipError_t code() const { return code_; }
  const std::string& message() const { return message_; }

  static CudaStatus OK() { return CudaStatus(); }
  static CudaStatus OkStatus() { return CudaStatus(); }

  // Keeps the first error (sticky), as the pano classes rely on.
  void Update(const CudaStatus& other) {
    if (ok() && !other.ok()) *this = other;
  }

  friend std::ostream& operator<<(std::ostream& os, const CudaStatus& s) {
    return os << "Code: " << static_cast<int>(s.code_) << ", Message: " << s.message_;
  }

 private:
  hipError_t code_ = hipSuccess;
  std::string message_ = "OK";
};

template <typename T>
class CudaStatusOr {
 public:
  CudaStatusOr(const T& value) : value_(value) {}
  CudaStatusOr(T&& value) : value_(std::move(value)) {}
  CudaStatusOr(const CudaStatus& status) : status_(status) { require_error(); }
  CudaStatusOr(CudaStatus&& status) : status_(std::move(status)) { require_error(); }

  bool ok() const { return status_.ok(); }
  const CudaStatus& status() const { return status_; }

  const T& ValueOrDie() const {
    check();
    return *value_;
  }
  T& ValueOrDie() {
    check();
    return *value_;
  }
  T ConsumeValueOrDie() {
    check();
    T v = std::move(*value_);
    value_.reset();
    return v;
  }

 private:
  void require_error() const {
    if (status_.ok()) throw std::invalid_argument("CudaStatusOr: an OK status needs a value");
  }
  void check() const {
    if (!ok() || !value_) throw std::runtime_error("CudaStatusOr: no value: " + status_.message());
  }
  CudaStatus status_;
  std::optional<T> value_;
};

#define CUDA_RETURN_IF_ERROR(expr)              \
  do {                                          \
    const CudaStatus _hm_status((expr));        \
    if (!_hm_status.ok()) return _hm_status;    \
  } while (0)
