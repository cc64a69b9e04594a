// device_memory.h — HEAL device tensors in MI355X HBM.
//
// Same public contract as the reference's include/device_memory.h:11-39:
// DeviceTensor<T> with reshape/print/print_metadata, and
// allocate_on_hardware / host_to_device / device_to_host in namespace
// lattica_hw_api, instantiated for int32_t, int64_t and double.
//
// A DeviceTensor is host metadata (dims, strides in elements) over a
// refcounted device buffer; views made by expand/squeeze/unsqueeze/reshape
// share the buffer.  Every buffer carries a unique id and a write version,
// bumped by every operation that writes into it, so derived device data (NTT
// plans, validated permutation tables) can be cached safely.
#ifndef HEAL_AMD_DEVICE_MEMORY_H
#define HEAL_AMD_DEVICE_MEMORY_H

#include <torch/torch.h>

#include <cstdint>
#include <memory>
#include <vector>

namespace lattica_hw_api {
namespace detail {

// One device allocation.  Freed (stream-ordered) when the last view drops it.
struct Buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    uint64_t id = 0;
    uint64_t version = 0;
    ~Buffer();
};

std::shared_ptr<Buffer> new_buffer(size_t bytes, bool zero);

}  // namespace detail
}  // namespace lattica_hw_api

template <typename T>
struct DeviceTensor {
    std::vector<int64_t> dims;
    std::vector<int64_t> strides;  // in elements; 0 along broadcast axes
    std::shared_ptr<lattica_hw_api::detail::Buffer> buffer;
    int64_t offset = 0;  // elements from the buffer start (views made by get_slice)

    // reference semantics (example_impl/device_memory_impl.cpp:42-84): the new
    // shape must have as many elements as the dims with non-zero stride; the
    // tensor gets C-contiguous strides for the new shape
    void reshape(const std::vector<int64_t>& new_dims);
    void print() const;
    void print_metadata() const;
    // C-contiguous, ignoring size-1 dims
    bool is_contiguous() const;
    T* data_ptr() const { return buffer ? static_cast<T*>(buffer->ptr) + offset : nullptr; }
};

namespace lattica_hw_api {

template <typename T>
std::shared_ptr<DeviceTensor<T>> allocate_on_hardware(const std::vector<int64_t>& dims);

template <typename T>
std::shared_ptr<DeviceTensor<T>> host_to_device(const torch::Tensor& tensor);

template <typename T>
torch::Tensor device_to_host(const std::shared_ptr<DeviceTensor<T>>& memory);

}  // namespace lattica_hw_api

#endif  // HEAL_AMD_DEVICE_MEMORY_H
