// The part of the VapourSynth API 3 (API version 0x30002) that the svp1 / svp2
// plugins use, declared by layout: the plugins read the VSAPI function table at the
// byte offsets the reference binds (crates/svpflow1/src/vs.rs:48-65,
// crates/svpflow2/src/vs.rs:53-73, analyse_filter.rs:194), and the VSVideoInfo /
// VSFormat fields at the offsets it reads (vs.rs:71-81, video_format.rs:3, 40-51).
// Every object the core hands out is opaque.
#pragma once
#include <cstddef>
#include <cstdint>

namespace vs3 {

struct Core;
struct Map;
struct Node;
struct FrameRef;
struct FrameContext;
struct Plugin;
struct Api;  // the VSAPI table: only read at the offsets below

// VSFormat: the name, then the preset id at byte 32 and the sample layout.
struct Format {
    char name[32];
    int32_t id;
    int32_t color_family;
    int32_t sample_type;
    int32_t bits_per_sample;
    int32_t bytes_per_sample;
    int32_t sub_sampling_w;
    int32_t sub_sampling_h;
    int32_t num_planes;
};
static_assert(offsetof(Format, id) == 32, "VSFormat id");

struct VideoInfo {
    const Format* format;
    int64_t fps_num, fps_den;
    int32_t width, height, num_frames, flags;
};
static_assert(sizeof(VideoInfo) == 40, "VSVideoInfo");

constexpr int32_t kApiVersion = 0x30002;
constexpr int32_t kPresetYUV420P8 = 3000010;
constexpr int32_t kPresetGray8 = 1000010;
constexpr int32_t kFmParallel = 100;
constexpr int kArInitial = 0, kArAllFramesReady = 2;
constexpr int kPaReplace = 0;

// Calling conventions of the callbacks the plugins hand to the core.
using PublicFunction = void (*)(const Map* in, Map* out, void* user, Core* core, const Api* api);
using FilterInit = void (*)(Map* in, Map* out, void** instance, Node* node, Core* core, const Api* api);
using FilterGetFrame = const FrameRef* (*)(int n, int activation, void** instance, void** frame_data,
                                            FrameContext* ctx, Core* core, const Api* api);
using FilterFree = void (*)(void* instance, Core* core, const Api* api);
using ConfigPlugin = void (*)(const char* identifier, const char* ns, const char* name, int api_version,
                              int readonly, Plugin* plugin);
using RegisterFunction = void (*)(const char* name, const char* args, PublicFunction fn, void* user, Plugin* plugin);

// Entries of the VSAPI table (byte offset, function type).
namespace fn {
using FreeFrame = void (*)(const FrameRef*);
using FreeNode = void (*)(Node*);
using NewVideoFrame = FrameRef* (*)(const Format*, int width, int height, const FrameRef* prop_src, Core*);
using CreateFilter = void (*)(const Map* in, Map* out, const char* name, FilterInit, FilterGetFrame, FilterFree,
                              int mode, int flags, void* instance, Core*);
using SetError = void (*)(Map*, const char* message);
using GetFormatPreset = const Format* (*)(int id, Core*);
using GetFrameFilter = const FrameRef* (*)(int n, Node*, FrameContext*);
using RequestFrameFilter = void (*)(int n, Node*, FrameContext*);
using GetStride = int (*)(const FrameRef*, int plane);
using GetReadPtr = const uint8_t* (*)(const FrameRef*, int plane);
using GetWritePtr = uint8_t* (*)(FrameRef*, int plane);
using GetVideoInfo = const VideoInfo* (*)(Node*);
using SetVideoInfo = void (*)(const VideoInfo*, int outputs, Node*);
using GetFramePropsRO = const Map* (*)(const FrameRef*);
using GetFramePropsRW = Map* (*)(FrameRef*);
using PropGetInt = int64_t (*)(const Map*, const char* key, int index, int* error);
using PropGetFloat = double (*)(const Map*, const char* key, int index, int* error);
using PropGetData = const char* (*)(const Map*, const char* key, int index, int* error);
using PropGetDataSize = int (*)(const Map*, const char* key, int index, int* error);
using PropGetNode = Node* (*)(const Map*, const char* key, int index, int* error);
using PropSetInt = int (*)(Map*, const char* key, int64_t value, int append);
}  // namespace fn

enum Offset : size_t {
    kFreeFrame = 48,
    kFreeNode = 56,
    kNewVideoFrame = 72,
    kCreateFilter = 136,
    kSetError = 144,
    kGetFormatPreset = 176,
    kGetFrameFilter = 208,
    kRequestFrameFilter = 216,
    kGetStride = 240,
    kGetReadPtr = 248,
    kGetWritePtr = 256,
    kGetVideoInfo = 304,
    kSetVideoInfo = 312,
    kGetFramePropsRO = 344,
    kGetFramePropsRW = 352,
    kPropGetInt = 392,
    kPropGetFloat = 400,
    kPropGetData = 408,
    kPropGetDataSize = 416,
    kPropGetNode = 424,
    kPropSetInt = 456,
};

// The table entry at `off` (nullptr if the table has none).
template <class F>
inline F entry(const Api* api, size_t off) {
    F f = nullptr;
    if (api) f = *reinterpret_cast<const F*>(reinterpret_cast<const uint8_t*>(api) + off);
    return f;
}

// The typed accessors the filters use.
struct Table {
    const Api* api;
    fn::FreeFrame free_frame() const { return entry<fn::FreeFrame>(api, kFreeFrame); }
    fn::FreeNode free_node() const { return entry<fn::FreeNode>(api, kFreeNode); }
    fn::NewVideoFrame new_video_frame() const { return entry<fn::NewVideoFrame>(api, kNewVideoFrame); }
    fn::CreateFilter create_filter() const { return entry<fn::CreateFilter>(api, kCreateFilter); }
    fn::SetError set_error() const { return entry<fn::SetError>(api, kSetError); }
    fn::GetFormatPreset get_format_preset() const { return entry<fn::GetFormatPreset>(api, kGetFormatPreset); }
    fn::GetFrameFilter get_frame_filter() const { return entry<fn::GetFrameFilter>(api, kGetFrameFilter); }
    fn::RequestFrameFilter request_frame_filter() const {
        return entry<fn::RequestFrameFilter>(api, kRequestFrameFilter);
    }
    fn::GetStride get_stride() const { return entry<fn::GetStride>(api, kGetStride); }
    fn::GetReadPtr get_read_ptr() const { return entry<fn::GetReadPtr>(api, kGetReadPtr); }
    fn::GetWritePtr get_write_ptr() const { return entry<fn::GetWritePtr>(api, kGetWritePtr); }
    fn::GetVideoInfo get_video_info() const { return entry<fn::GetVideoInfo>(api, kGetVideoInfo); }
    fn::SetVideoInfo set_video_info() const { return entry<fn::SetVideoInfo>(api, kSetVideoInfo); }
    fn::GetFramePropsRO get_frame_props_ro() const { return entry<fn::GetFramePropsRO>(api, kGetFramePropsRO); }
    fn::GetFramePropsRW get_frame_props_rw() const { return entry<fn::GetFramePropsRW>(api, kGetFramePropsRW); }
    fn::PropGetInt prop_get_int() const { return entry<fn::PropGetInt>(api, kPropGetInt); }
    fn::PropGetFloat prop_get_float() const { return entry<fn::PropGetFloat>(api, kPropGetFloat); }
    fn::PropGetData prop_get_data() const { return entry<fn::PropGetData>(api, kPropGetData); }
    fn::PropGetDataSize prop_get_data_size() const { return entry<fn::PropGetDataSize>(api, kPropGetDataSize); }
    fn::PropGetNode prop_get_node() const { return entry<fn::PropGetNode>(api, kPropGetNode); }
    fn::PropSetInt prop_set_int() const { return entry<fn::PropSetInt>(api, kPropSetInt); }
};

}  // namespace vs3
