// DLPack v0.8 ABI (dmlc/dlpack, include/dlpack/dlpack.h), restated: the
// structs the engine exports and the binding wraps in "dltensor" capsules.
#pragma once
#include <stdint.h>

namespace gm {
struct DLDevice {
  int32_t device_type; // kDLROCM = 10
  int32_t device_id;
};
struct DLDataType {
  uint8_t code; // 0 int, 1 uint, 2 float
  uint8_t bits;
  uint16_t lanes;
};
struct DLTensor {
  void *data;
  DLDevice device;
  int32_t ndim;
  DLDataType dtype;
  int64_t *shape;
  int64_t *strides;
  uint64_t byte_offset;
};
struct DLManagedTensor {
  DLTensor dl_tensor;
  void *manager_ctx;
  void (*deleter)(DLManagedTensor *self);
};
} // namespace gm
