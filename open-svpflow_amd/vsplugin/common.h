// Staging of VapourSynth frames to and from the device for the svp1 / svp2 plugin
// shells: a frame's planes are copied whole (stride x rows), so the device copy keeps
// the host frame's strides and every padding byte.
#pragma once
#include <cstring>
#include <string>

#include "svpf.h"
#include "vsapi3.h"

namespace svpvs {

#define SVPVS_EXPORT extern "C" __attribute__((visibility("default")))

inline int plane_rows(int rows, int plane) { return plane == 0 ? rows : rows / 2; }

// Device copy of a 3-plane 4:2:0 frame (a source frame or a Super frame).
struct DevFrame {
    uint8_t* p[3] = {nullptr, nullptr, nullptr};
    int stride[3] = {0, 0, 0};
    int rows = 0;

    // Holds planes of these strides and `luma_rows` rows (allocated on first use).
    bool fit(svpf_ctx* ctx, const int strides[3], int luma_rows) {
        if (p[0] && rows == luma_rows && stride[0] == strides[0] && stride[1] == strides[1] &&
            stride[2] == strides[2])
            return true;
        release(ctx);
        for (int k = 0; k < 3; k++) {
            void* d = nullptr;
            if (svpf_device_alloc(ctx, (size_t)strides[k] * (size_t)plane_rows(luma_rows, k), &d) != SVPF_OK) {
                release(ctx);
                return false;
            }
            p[k] = (uint8_t*)d;
            stride[k] = strides[k];
        }
        rows = luma_rows;
        return true;
    }
    void release(svpf_ctx* ctx) {
        for (int k = 0; k < 3; k++) {
            if (p[k]) svpf_device_free(ctx, p[k]);
            p[k] = nullptr;
        }
        rows = 0;
    }
    svpf_frame frame() const { return svpf_frame{p[0], p[1], p[2], stride[0], stride[1], stride[2]}; }
};

inline void read_strides(const vs3::Table& t, const vs3::FrameRef* f, int out[3]) {
    const auto get_stride = t.get_stride();
    for (int k = 0; k < 3; k++) out[k] = get_stride(f, k);
}

// Host frame -> device (planes of `luma_rows` rows).
inline bool upload(svpf_ctx* ctx, DevFrame& d, const vs3::Table& t, const vs3::FrameRef* f, int luma_rows) {
    int s[3];
    read_strides(t, f, s);
    if (!d.fit(ctx, s, luma_rows)) return false;
    const auto read = t.get_read_ptr();
    for (int k = 0; k < 3; k++) {
        const uint8_t* h = read(f, k);
        if (!h || svpf_copy_h2d(ctx, d.p[k], h, (size_t)s[k] * (size_t)plane_rows(luma_rows, k)) != SVPF_OK)
            return false;
    }
    return true;
}

// Device -> host frame of the same strides.
inline bool download(svpf_ctx* ctx, const DevFrame& d, const vs3::Table& t, vs3::FrameRef* f) {
    const auto write = t.get_write_ptr();
    for (int k = 0; k < 3; k++) {
        uint8_t* h = write(f, k);
        if (!h || t.get_stride()(f, k) != d.stride[k] ||
            svpf_copy_d2h(ctx, h, d.p[k], (size_t)d.stride[k] * (size_t)plane_rows(d.rows, k)) != SVPF_OK)
            return false;
    }
    return true;
}

// The bytes of a data argument ("opt"); empty when absent.
inline std::string read_data(const vs3::Table& t, const vs3::Map* in, const char* key) {
    const auto get_data = t.prop_get_data();
    const auto get_size = t.prop_get_data_size();
    if (!get_data || !get_size) return {};
    int err = 0;
    const char* p = get_data(in, key, 0, &err);
    if (!p || err) return {};
    const int n = get_size(in, key, 0, &err);
    if (n <= 0 || err) return {};
    return std::string(p, (size_t)n);
}

// 8-bit 4:2:0 YUV with three planes (the device path's only layout).
inline bool is_yuv420p8(const vs3::Format* f) {
    return f && f->bytes_per_sample == 1 && f->bits_per_sample == 8 && f->sub_sampling_w == 1 &&
           f->sub_sampling_h == 1 && f->num_planes == 3;
}

inline void free_node(const vs3::Table& t, vs3::Node* n) {
    if (n && t.free_node()) t.free_node()(n);
}

}  // namespace svpvs
