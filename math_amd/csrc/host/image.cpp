// C ABI of the view's image path (src/gpu/dft/compute.cpp:91-200,
// DftImage): luminance of an RGBA rectangle, forward 2-D DFT, |X|/(w*h) into
// an R32F image.
#include "capi_util.hpp"

#include <algorithm>
#include <string>

using mdft::Context;
using mdft::Error;
using mdft::ImageIO;
using mdft::LinePlan;

struct mdft_image_s {
    int device = 0;
    uint32_t w = 0, h = 0, fmt = 0;
    std::shared_ptr<const LinePlan> row, col;
    float2* buf = nullptr;  // complex image between the passes (the reference's x_d)
    float2* scratch = nullptr;
    uint64_t scratch_elems = 0;
    mdft::ScratchOrder order;  // calls sharing buf and scratch on different streams
};

namespace {

void launch(hipError_t e, const char* what) { mdft::hip_check(e, what); }

}  // namespace

extern "C" {

int mdft_image_create(mdft_image* image, uint32_t width, uint32_t height, uint32_t pixel_format)
{
    MDFT_API_BEGIN
    if (!image) throw Error(MDFT_ERR_INVALID, "mdft_image_create: image is NULL");
    *image = nullptr;
    if (pixel_format != MDFT_PIXEL_RGBA32F && pixel_format != MDFT_PIXEL_RGBA8)
        throw Error(MDFT_ERR_INVALID, "mdft_image_create: unknown pixel format " + std::to_string(pixel_format));
    mdft::compute_m((long long)width);
    mdft::compute_m((long long)height);
    Context& ctx = Context::get(Context::current_device());
    auto p = std::make_unique<mdft_image_s>();
    p->device = ctx.device();
    p->w = width;
    p->h = height;
    p->fmt = pixel_format;
    p->row = ctx.line_plan((int)width, false);
    p->col = ctx.line_plan((int)height, true);
    p->scratch_elems = std::max(width > 1 ? p->row->scratch_elems((int)height) : uint64_t(0),
                                height > 1 ? p->col->scratch_elems((int)width) : uint64_t(0));
    mdft::DeviceGuard g(p->device);
    mdft::hip_check(hipMalloc(&p->buf, sizeof(float2) * (size_t)width * height), "hipMalloc image buffer");
    if (p->scratch_elems)
        mdft::hip_check(hipMalloc(&p->scratch, sizeof(float2) * p->scratch_elems), "hipMalloc scratch");
    *image = p.release();
    MDFT_API_END
}

int mdft_image_destroy(mdft_image image)
{
    MDFT_API_BEGIN
    if (!image) return MDFT_OK;
    {
        mdft::DeviceGuard g(image->device);
        (void)hipDeviceSynchronize();
        if (image->buf) (void)hipFree(image->buf);
        if (image->scratch) (void)hipFree(image->scratch);
        image->order.release();
    }
    delete image;
    MDFT_API_END
}

int mdft_image_exec_device(mdft_image image, const void* src, uint32_t src_width, uint32_t src_height,
                           uint64_t src_pitch, uint32_t x, uint32_t y, float* dst, uint64_t dst_pitch, void* stream)
{
    MDFT_API_BEGIN
    if (!image) throw Error(MDFT_ERR_NO_BUFFERS, "No DFT buffers");
    if (!src || !dst) throw Error(MDFT_ERR_DATA_SIZE, "image pointer is NULL");
    // the rectangle lies inside the source image (compute.cpp:113-118)
    if ((uint64_t)x + image->w > src_width || (uint64_t)y + image->h > src_height || src_pitch < src_width ||
        dst_pitch < image->w)
        throw Error(MDFT_ERR_INVALID, "image rectangle (" + std::to_string(x) + ", " + std::to_string(y) + ", " +
                                          std::to_string(image->w) + " x " + std::to_string(image->h) +
                                          ") is not inside the " + std::to_string(src_width) + " x " +
                                          std::to_string(src_height) + " source, or a pitch is too small");
    mdft::DeviceGuard g(image->device);
    Context& ctx = Context::get(image->device);
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const int w = (int)image->w, h = (int)image->h;
    const size_t psize = image->fmt == MDFT_PIXEL_RGBA8 ? 4 : 16;
    ImageIO io{};
    io.img = static_cast<const char*>(src) + psize * ((uint64_t)y * src_pitch + x);
    io.fmt = image->fmt == MDFT_PIXEL_RGBA8 ? mdft::PIXEL_RGBA8 : mdft::PIXEL_RGBA32F;
    io.mag = dst;
    // TO_MUL = 1 / (w*h) in double, rounded to float (compute.cpp:128-129)
    io.mscale = (float)(1.0 / ((double)w * (double)h));
    const LinePlan& row = *image->row;
    const LinePlan& col = *image->col;
    const bool fused = !row.large() && !col.large() && (w > 1 || h > 1);
    image->order.before(s);
    if (!fused) {
        // luminance -> complex buffer, ordinary transform in place, magnitude
        io.iel = 1, io.ils = (int64_t)src_pitch;
        io.mel = 1, io.mls = (int64_t)dst_pitch;
        launch(mdft::launch_image_luminance(io, image->buf, w, h, s), "image luminance launch");
        ctx.enqueue_transform(row, col, w, h, false, image->buf, image->buf, image->scratch, s);
        launch(mdft::launch_image_magnitude(io, image->buf, w, h, s), "image magnitude launch");
        image->order.after(s);
        return MDFT_OK;
    }
    if (w > 1) {
        // rows: luminance in; magnitude out as well when there is no column pass
        io.iel = 1, io.ils = (int64_t)src_pitch;
        io.mel = 1, io.mls = (int64_t)dst_pitch;
        const mdft::AxisJob j = ctx.make_job(row, nullptr, image->buf, 1, w, h, false);
        const int mode = mdft::IO_LUM | (h == 1 ? mdft::IO_MAG : 0);
        launch(mdft::launch_image_axis(row.shape, j, io, row.blue, false, mode, false, s), "image row pass launch");
    }
    if (h > 1) {
        // columns: magnitude out; luminance in as well when there was no row pass
        io.iel = (int64_t)src_pitch, io.ils = 1;
        io.mel = (int64_t)dst_pitch, io.mls = 1;
        const mdft::AxisJob j = ctx.make_job(col, image->buf, nullptr, w, 1, w, false);
        const int mode = mdft::IO_MAG | (w == 1 ? mdft::IO_LUM : 0);
        const bool wide = ctx.wide_tiles(col.shape, col.blue, true, w);
        launch(mdft::launch_image_axis(col.shape, j, io, col.blue, true, mode, wide, s), "image column pass launch");
    }
    image->order.after(s);
    MDFT_API_END
}

}  // extern "C"
