// Device discovery and selection (the reference's include/gpu_devices.hpp:
// 9-23, same names).  This build drives one GPU per search: a selection of
// more than one device is rejected.
#ifndef RATPOINTS_GPU_GPU_DEVICES_HPP
#define RATPOINTS_GPU_GPU_DEVICES_HPP

#include <string>
#include <vector>

namespace ratpoints_gpu {

struct GpuDevice {
    int id;
    std::string name;
    int major;
    int minor;
    int multiprocessors;
    unsigned long long memory_bytes;
    std::string pci_bus_id;
};

// "all" gives an empty list, meaning every visible device.
std::vector<int> parse_device_selection(const std::string &text);
std::vector<GpuDevice> available_gpu_devices();
std::vector<GpuDevice> select_gpu_devices(const std::vector<int> &ids);
// Binds the calling thread's searches to device `id`.
void activate_gpu(int id);
int active_gpu();

}  // namespace ratpoints_gpu

#endif
