// Device discovery over the C ABI (rps_device_count / rps_device_info).
// The --devices syntax and its messages are the reference's interface
// (src/gpu_devices.cpp:21-53); a search uses exactly one device.
#include "gpu_devices.hpp"

#include <algorithm>
#include <limits>
#include <stdexcept>

#include "ratpoints_sieve.h"

namespace ratpoints_gpu {
namespace {

thread_local int bound_device = 0;

}  // namespace

namespace {

// The comma-separated fields of a device list ("" gives one empty field).
std::vector<std::string> comma_fields(const std::string &text) {
    std::vector<std::string> fields(1);
    for (char c : text) {
        if (c == ',') {
            fields.emplace_back();
        } else {
            fields.back() += c;
        }
    }
    return fields;
}

// A device id: decimal digits only, within int.
bool device_id(const std::string &field, int &id) {
    long long v = 0;
    for (char c : field) {
        if (c < '0' || c > '9') {
            return false;
        }
        v = 10 * v + (c - '0');
        if (v > std::numeric_limits<int>::max()) {
            return false;
        }
    }
    id = static_cast<int>(v);
    return true;
}

}  // namespace

std::vector<int> parse_device_selection(const std::string &text) {
    if (text == "all") {
        return {};
    }
    const std::vector<std::string> fields = comma_fields(text);
    std::vector<int> ids;
    for (size_t k = 0; k < fields.size(); ++k) {
        if (fields[k].empty()) {
            // A trailing empty field means a trailing comma (or no list).
            throw std::invalid_argument(k + 1 == fields.size()
                ? "--devices requires all or a nonempty GPU list"
                : "empty GPU ID in --devices");
        }
        int id = 0;
        if (!device_id(fields[k], id)) {
            throw std::invalid_argument(
                "--devices requires all or comma-separated nonnegative GPU IDs");
        }
        if (std::count(ids.begin(), ids.end(), id)) {
            throw std::invalid_argument("duplicate GPU ID in --devices");
        }
        ids.push_back(id);
    }
    return ids;
}

std::vector<GpuDevice> available_gpu_devices() {
    int count = 0;
    if (rps_device_count(&count) != RPS_OK) {
        throw std::runtime_error(rps_last_error(nullptr));
    }
    if (count == 0) {
        throw std::runtime_error(
            "no visible HIP devices; check the driver and HIP_VISIBLE_DEVICES");
    }
    std::vector<GpuDevice> devices;
    for (int id = 0; id < count; ++id) {
        char name[256] = {0};
        char pci[64] = {0};
        int units = 0, major = 0, minor = 0;
        uint64_t memory = 0;
        if (rps_device_info(id, name, sizeof(name), pci, sizeof(pci), &units,
                            &memory, &major, &minor) != RPS_OK) {
            throw std::runtime_error(rps_last_error(nullptr));
        }
        devices.push_back({id, name, major, minor, units,
                           static_cast<unsigned long long>(memory), pci});
    }
    return devices;
}

std::vector<GpuDevice> select_gpu_devices(const std::vector<int> &ids) {
    std::vector<GpuDevice> visible = available_gpu_devices();
    if (ids.empty()) {
        return {visible.front()};
    }
    if (ids.size() > 1) {
        throw std::invalid_argument(
            "this backend searches on one GPU; select it with --devices N");
    }
    const int id = ids.front();
    if (id < 0 || static_cast<size_t>(id) >= visible.size()) {
        throw std::invalid_argument("GPU ID " + std::to_string(id)
            + " is not visible; run --list-devices (IDs are HIP-visible ordinals)");
    }
    return {visible[static_cast<size_t>(id)]};
}

void activate_gpu(int id) {
    int count = 0;
    if (rps_device_count(&count) != RPS_OK) {
        throw std::runtime_error(rps_last_error(nullptr));
    }
    if (id < 0 || id >= count) {
        throw std::runtime_error("hipSetDevice: invalid device ordinal");
    }
    bound_device = id;
}

int active_gpu() { return bound_device; }

}  // namespace ratpoints_gpu
