/* artful-hip — the differentiable analytic-disc renderer (include/artful/adjoint.h): frames from a
 * description and a camera, the trace and adjoint kernels on one stream, the workgroups' partial
 * gradients summed in double in a fixed order, and the chain from the frame's fields to the spin and
 * the eye's position by central differences of the FP64 frame (a smooth function of both). */
#include <stdio.h>

#include "../scene/scene_internal.h"
#include "artful/adjoint.h"
#include "artful_hip.h"

struct ArtfHoleAdjoint {
    int width, height, side;
    void* stream;
    void* events[4];
    float* map;        /* device: the map's layout, only its colour table filled */
    float* hdr[3];     /* device, width*height */
    float* rays[3];    /* device, width*height*side^2 when side > 1 */
    float* seed[3];    /* device, width*height */
    double* partials;  /* device */
    double* partialsHost;
    int blocks;
    double forwardMs, adjointMs;
};

static void freePlanes(float** p) {
    for (int c = 0; c < 3; ++c) artfHipFree(p[c]), p[c] = NULL;
}

static bool allocPlanes(float** p, size_t n) {
    for (int c = 0; c < 3; ++c)
        if (!(p[c] = artfHipMalloc(n * sizeof(float)))) return false;
    return true;
}

ArtfHoleAdjoint* artfHoleAdjointNew(int width, int height, int device) {
    if (width < 1 || height < 1 || device < 0 || device >= artfHipDeviceCount() || artfHipSetDevice(device)) {
        artfLog("artful: no HIP device %d for the adjoint", device);
        return NULL;
    }
    ArtfHoleAdjoint* a = ARTF_NEW(ArtfHoleAdjoint);
    a->width = width, a->height = height;
    a->stream = artfHipStreamCreate();
    bool ready = a->stream != NULL;
    for (int i = 0; i < 4; ++i) ready = ready && (a->events[i] = artfHipEventCreate()) != NULL;
    const size_t pixels = (size_t)width * height;
    ready = ready && (a->map = artfHipMalloc((size_t)ARTF_HOLE_MAP_FLOATS * sizeof(float))) != NULL;
    ready = ready && allocPlanes(a->hdr, pixels) && allocPlanes(a->seed, pixels);
    if (ready) {
        float colors[ARTF_HOLE_COLORS * 3];
        artfDiscColors(colors);
        ready = !artfHipMemset(a->map, 0, (size_t)ARTF_HOLE_MAP_FLOATS * sizeof(float), a->stream) &&
                !artfHipUpload(a->map + ARTF_HOLE_COLORS_AT, colors, sizeof colors, a->stream) && !artfHipStreamSync(a->stream);
    }
    if (!ready) {
        artfLog("artful: the adjoint could not set up its device memory");
        artfHoleAdjointFree(a);
        return NULL;
    }
    return a;
}

void artfHoleAdjointFree(ArtfHoleAdjoint* a) {
    if (!a) return;
    if (a->stream) artfHipStreamSync(a->stream);
    freePlanes(a->hdr), freePlanes(a->rays), freePlanes(a->seed);
    artfHipFree(a->map);
    artfHipFree(a->partials);
    artfFree(a->partialsHost);
    for (int i = 0; i < 4; ++i) artfHipEventDestroy(a->events[i]);
    artfHipStreamDestroy(a->stream);
    artfFree(a);
}

/* The trace arguments for this description and camera; false when they are not differentiable here. */
static bool prepare(ArtfHoleAdjoint* a, const ArtfBlackHoleDesc* desc, const ArtfCamera* camera, const ArtfHoleTracing* t, ArtfHipTraceArgs* args) {
    const ArtfBlackHoleDesc d = artfHoleClamped(desc);
    if (d.stars > 0 || artfVec3Length(d.velocity) > 0 || desc->discInner <= 0) {
        artfLog("artful: the adjoint needs stars 0, the eye at rest and an explicit discInner");
        return false;
    }
    memset(args, 0, sizeof *args);
    uint32_t bloomPixels = 0;
    artfHoleFrameOf(&d, camera, 0.0, a->width, a->height, 1, &args->frame, &bloomPixels);
    artfHoleFrame64Of(&d, camera, a->width, a->height, &args->frame64);
    const int side = d.samples;
    if (side != a->side) {
        freePlanes(a->rays);
        a->side = 0;
        if (side > 1 && !allocPlanes(a->rays, (size_t)a->width * a->height * side * side)) return false;
        a->side = side;
    }
    args->map = a->map;
    args->fluid = a->map; /* (not read by the analytic disc) */
    float** out = side > 1 ? a->rays : a->hdr;
    for (int c = 0; c < 3; ++c) args->out[c] = out[c];
    args->precision = t ? t->precision : ARTF_PRECISION_FP32;
    args->maxSteps = t ? t->maxSteps : 0;
    args->escapeRadius = t ? t->escapeRadius : 0.f;
    args->disc = ARTF_DISC_ANALYTIC;
    args->emissivity = t ? t->emissivity : 3.f;
    args->jitter = t ? t->jitter : 0;
    return true;
}

static bool readHdr(ArtfHoleAdjoint* a, const ArtfHoleFrame* f, float* hdr) {
    if (a->side > 1 && artfHipResolve(f, a->rays, a->hdr, a->stream)) return false;
    if (!hdr) return !artfHipStreamSync(a->stream);
    const size_t pixels = (size_t)a->width * a->height;
    for (int c = 0; c < 3; ++c)
        if (artfHipDownload(hdr + c * pixels, a->hdr[c], pixels * sizeof(float), a->stream)) return false;
    return true;
}

bool artfHoleAdjointRender(ArtfHoleAdjoint* a, const ArtfBlackHoleDesc* desc, const ArtfCamera* camera, const ArtfHoleTracing* t, float* hdr) {
    ArtfHipTraceArgs args;
    if (!prepare(a, desc, camera, t, &args)) return false;
    if (artfHipEventRecord(a->events[0], a->stream) || artfHipTrace(&args, a->stream) || artfHipEventRecord(a->events[1], a->stream)) return false;
    if (!readHdr(a, &args.frame, hdr)) return false;
    a->forwardMs = artfHipEventMs(a->events[0], a->events[1]);
    return true;
}

/* The 26 frame fields the rays read, in the gradient's order. */
static void fieldsOf(const ArtfHoleFrame64* f, double* out) {
    const double* groups[8] = {f->forward, f->right, f->up, f->metric, f->place, f->er, f->etheta, f->ephi};
    const int counts[8] = {3, 3, 3, 4, 4, 3, 3, 3};
    int k = 0;
    for (int g = 0; g < 8; ++g)
        for (int i = 0; i < counts[g]; ++i) out[k++] = groups[g][i];
}

bool artfHoleAdjointGradient(ArtfHoleAdjoint* a, const ArtfBlackHoleDesc* desc, const ArtfCamera* camera, const ArtfHoleTracing* t, const float* seed,
                             float* hdr, ArtfHoleGradient* out) {
    memset(out, 0, sizeof *out);
    ArtfHipTraceArgs trace;
    if (!prepare(a, desc, camera, t, &trace)) return false;
    if (trace.precision != ARTF_PRECISION_FP64) trace.precision = ARTF_PRECISION_FP32;
    const int blocks = artfHipAdjointBlocks(&trace.frame);
    if (blocks > a->blocks) {
        artfHipFree(a->partials);
        artfFree(a->partialsHost);
        a->partials = artfHipMalloc((size_t)blocks * 32 * sizeof(double));
        a->partialsHost = artfMalloc((size_t)blocks * 32 * sizeof(double));
        a->blocks = a->partials ? blocks : 0;
        if (!a->partials) return false;
    }
    const size_t pixels = (size_t)a->width * a->height;
    for (int c = 0; c < 3; ++c)
        if (artfHipUpload(a->seed[c], seed + c * pixels, pixels * sizeof(float), a->stream)) return false;
    ArtfHipAdjointArgs args;
    memset(&args, 0, sizeof args);
    args.trace = trace;
    for (int c = 0; c < 3; ++c) args.seed[c] = a->seed[c];
    args.partials = a->partials;
    if (artfHipEventRecord(a->events[2], a->stream) || artfHipAdjoint(&args, a->stream) || artfHipEventRecord(a->events[3], a->stream)) return false;
    if (artfHipDownload(a->partialsHost, a->partials, (size_t)blocks * 32 * sizeof(double), a->stream)) return false;
    a->adjointMs = artfHipEventMs(a->events[2], a->events[3]);
    if (!readHdr(a, &trace.frame, hdr)) return false;
    double sum[32] = {0};
    for (int b = 0; b < blocks; ++b)
        for (int k = 0; k < 32; ++k) sum[k] += a->partialsHost[(size_t)b * 32 + k];
    for (int k = 0; k < 26; ++k) out->frame[k] = sum[k];
    out->spinDirect = sum[26];
    out->discInner = sum[27], out->discOuter = sum[28], out->emissivity = sum[29], out->gain = sum[30];

    /* the frame's fields as functions of the spin and the eye, by central differences in doubles */
    const ArtfBlackHoleDesc d = artfHoleClamped(desc);
    const double eye[3] = {camera->position.x, camera->position.y, camera->position.z};
    const double reach = sqrt(eye[0] * eye[0] + eye[1] * eye[1] + eye[2] * eye[2]);
    for (int p = 0; p < 4; ++p) {
        const double h = p == 0 ? 1e-6 : 1e-6 * (reach > 1 ? reach : 1);
        double lo[3] = {eye[0], eye[1], eye[2]}, hi[3] = {eye[0], eye[1], eye[2]};
        double spinLo = d.spin, spinHi = d.spin;
        if (p == 0) spinLo -= h, spinHi += h;
        else lo[p - 1] -= h, hi[p - 1] += h;
        ArtfHoleFrame64 fl, fh;
        artfHoleFrame64At(&d, camera, spinLo, lo, a->width, a->height, &fl);
        artfHoleFrame64At(&d, camera, spinHi, hi, a->width, a->height, &fh);
        double vl[26], vh[26], dot = 0;
        fieldsOf(&fl, vl);
        fieldsOf(&fh, vh);
        for (int k = 0; k < 26; ++k) dot += out->frame[k] * (vh[k] - vl[k]) / (2 * h);
        if (p == 0) out->spin = dot + out->spinDirect;
        else out->position[p - 1] = dot;
    }
    return true;
}

void artfHoleAdjointTimes(const ArtfHoleAdjoint* a, double* forwardMs, double* adjointMs) {
    if (forwardMs) *forwardMs = a->forwardMs;
    if (adjointMs) *adjointMs = a->adjointMs;
}
