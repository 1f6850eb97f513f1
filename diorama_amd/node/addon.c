/*
 * N-API addon over the C ABI (include/diorama.h), written in plain C against
 * node_api.h. Handles travel to JavaScript as externals; a non-zero status from
 * the library becomes a thrown Error carrying dio_last_error(). The JS classes in
 * index.js give these functions the method names of the reference's TypeScript
 * operators.
 */
#include <node_api.h>
#include <stdint.h>
#include <stdlib.h>
#include <stdio.h>
#include <string.h>

#include "diorama.h"

#define MAX_ARGS 10

typedef struct {
  napi_env env;
  size_t argc;
  napi_value argv[MAX_ARGS];
} call_t;

static int get_args(napi_env env, napi_callback_info info, call_t* c) {
  c->env = env;
  c->argc = MAX_ARGS;
  return napi_get_cb_info(env, info, &c->argc, c->argv, NULL, NULL) == napi_ok;
}

static napi_value undefined(napi_env env) {
  napi_value v;
  napi_get_undefined(env, &v);
  return v;
}

static napi_value fail(napi_env env, const char* message) {
  napi_throw_error(env, NULL, message);
  return NULL;
}

/* A library status: throw with the library's message when it is not DIO_OK. */
static int ok(napi_env env, int status) {
  if (status != DIO_OK) {
    napi_throw_error(env, NULL, dio_last_error());
    return 0;
  }
  return 1;
}

static int is_nullish(call_t* c, size_t i) {
  napi_valuetype t;
  if (i >= c->argc) return 1;
  napi_typeof(c->env, c->argv[i], &t);
  return t == napi_null || t == napi_undefined;
}

static void* handle(call_t* c, size_t i) {
  void* p = NULL;
  if (is_nullish(c, i)) return NULL;
  napi_get_value_external(c->env, c->argv[i], &p);
  return p;
}

static uint32_t u32(call_t* c, size_t i) {
  uint32_t v = 0;
  if (i < c->argc) napi_get_value_uint32(c->env, c->argv[i], &v);
  return v;
}

static double f64(call_t* c, size_t i) {
  double v = 0;
  if (i < c->argc) napi_get_value_double(c->env, c->argv[i], &v);
  return v;
}

/* A byte size or offset: a finite, non-negative integer below 2^53. Casting any
   other double to size_t is undefined behaviour, and a negative one would wrap. */
static int size_arg(call_t* c, size_t i, size_t* out) {
  double v = f64(c, i);
  if (!(v >= 0 && v <= 9007199254740991.0) || v != (double)(uint64_t)v) {
    napi_throw_range_error(c->env, NULL, "byte offsets and sizes must be non-negative integers");
    return 0;
  }
  *out = (size_t)v;
  return 1;
}

static napi_value wrap(napi_env env, void* p) {
  napi_value v;
  if (!p) {
    napi_get_null(env, &v);
    return v;
  }
  napi_create_external(env, p, NULL, NULL, &v);
  return v;
}

static napi_value number(napi_env env, double x) {
  napi_value v;
  napi_create_double(env, x, &v);
  return v;
}

/* Data pointer and byte length of a TypedArray or ArrayBuffer argument. */
static int bytes_of(call_t* c, size_t i, void** data, size_t* length) {
  bool is_typed = false, is_buffer = false;
  if (i >= c->argc) return 0;
  napi_is_typedarray(c->env, c->argv[i], &is_typed);
  if (is_typed) {
    napi_typedarray_type type;
    size_t count, offset;
    napi_value ab;
    napi_get_typedarray_info(c->env, c->argv[i], &type, &count, data, &ab, &offset);
    size_t item = (type == napi_int8_array || type == napi_uint8_array || type == napi_uint8_clamped_array) ? 1
                  : (type == napi_int16_array || type == napi_uint16_array)                                ? 2
                  : (type == napi_float64_array)                                                            ? 8
                                                                                                            : 4;
    *length = count * item;
    return 1;
  }
  napi_is_arraybuffer(c->env, c->argv[i], &is_buffer);
  if (is_buffer) return napi_get_arraybuffer_info(c->env, c->argv[i], data, length) == napi_ok;
  return 0;
}

/* A float[n] from a JS array (or typed array). */
static int floats(call_t* c, napi_value arr, float* out, uint32_t n) {
  for (uint32_t k = 0; k < n; k++) {
    napi_value e;
    double d = 0;
    if (napi_get_element(c->env, arr, k, &e) != napi_ok) return 0;
    napi_get_value_double(c->env, e, &d);
    out[k] = (float)d;
  }
  return 1;
}

static napi_value property(napi_env env, napi_value obj, const char* name) {
  napi_value v;
  napi_get_named_property(env, obj, name, &v);
  return v;
}

/* CameraData {view, proj, viewport} (src/gaussianRenderer.ts:11-15). */
static int camera_of(call_t* c, napi_value obj, dio_camera* cam) {
  return floats(c, property(c->env, obj, "view"), cam->view, 16) &&
         floats(c, property(c->env, obj, "proj"), cam->proj, 16) &&
         floats(c, property(c->env, obj, "viewport"), cam->viewport, 2);
}

#define FN(name) static napi_value name(napi_env env, napi_callback_info info)
#define ARGS          \
  call_t c;           \
  get_args(env, info, &c)

/* ------------------------------------------------------------- runtime -- */

FN(ctx_create) {
  ARGS;
  dio_ctx* ctx = NULL;
  if (!ok(env, dio_ctx_create((int)u32(&c, 0), &ctx))) return NULL;
  return wrap(env, ctx);
}
FN(ctx_synchronize) {
  ARGS;
  if (!ok(env, dio_ctx_synchronize(handle(&c, 0)))) return NULL;
  return undefined(env);
}
FN(ctx_destroy) {
  ARGS;
  dio_ctx_destroy(handle(&c, 0));
  return undefined(env);
}
FN(version) {
  napi_value v;
  (void)info;
  napi_create_string_utf8(env, dio_version(), NAPI_AUTO_LENGTH, &v);
  return v;
}

FN(buf_create) {
  ARGS;
  dio_buf* b = NULL;
  size_t bytes;
  if (!size_arg(&c, 1, &bytes)) return NULL;
  if (!ok(env, dio_buf_create(handle(&c, 0), bytes, &b))) return NULL;
  return wrap(env, b);
}
FN(buf_upload) {
  ARGS;
  void* data;
  size_t length, offset;
  if (!size_arg(&c, 1, &offset)) return NULL;
  if (!bytes_of(&c, 2, &data, &length)) return fail(env, "upload needs a TypedArray or ArrayBuffer");
  if (!ok(env, dio_buf_upload(handle(&c, 0), offset, data, length))) return NULL;
  return undefined(env);
}
FN(buf_download) {
  ARGS;
  size_t bytes, offset;
  if (!size_arg(&c, 1, &offset) || !size_arg(&c, 2, &bytes)) return NULL;
  void* data = NULL;
  napi_value ab;
  if (!ok(env, dio_buf_download(handle(&c, 0), offset, NULL, 0))) return NULL; /* range check first */
  if (bytes > dio_buf_size(handle(&c, 0)) - offset) return fail(env, "download past the end of the buffer");
  napi_create_arraybuffer(env, bytes, &data, &ab);
  if (!ok(env, dio_buf_download(handle(&c, 0), offset, data, bytes))) return NULL;
  return ab;
}
FN(buf_copy) {
  ARGS;
  size_t dst_offset, src_offset, bytes;
  if (!size_arg(&c, 1, &dst_offset) || !size_arg(&c, 3, &src_offset) || !size_arg(&c, 4, &bytes)) return NULL;
  if (!ok(env, dio_buf_copy(handle(&c, 0), dst_offset, handle(&c, 2), src_offset, bytes))) return NULL;
  return undefined(env);
}
FN(ctx_timestamp) {
  ARGS;
  if (!ok(env, dio_ctx_timestamp(handle(&c, 0), u32(&c, 1)))) return NULL;
  return undefined(env);
}
FN(ctx_elapsed) {
  ARGS;
  double ms = 0;
  if (!ok(env, dio_ctx_elapsed(handle(&c, 0), u32(&c, 1), u32(&c, 2), &ms))) return NULL;
  return number(env, ms);
}
FN(buf_clear) {
  ARGS;
  if (!ok(env, dio_buf_clear(handle(&c, 0)))) return NULL;
  return undefined(env);
}
FN(buf_size) {
  ARGS;
  return number(env, (double)dio_buf_size(handle(&c, 0)));
}
FN(buf_destroy) {
  ARGS;
  dio_buf_destroy(handle(&c, 0));
  return undefined(env);
}

/* ----------------------------------------------------------- PrefixSum -- */

FN(prefix_sum_create) {
  ARGS;
  dio_prefix_sum* s = NULL;
  if (!ok(env, dio_prefix_sum_create(handle(&c, 0), handle(&c, 1), u32(&c, 2), handle(&c, 3), &s))) return NULL;
  return wrap(env, s);
}
FN(prefix_sum_total) {
  ARGS;
  return wrap(env, dio_prefix_sum_total(handle(&c, 0)));
}
FN(prefix_sum_encode) {
  ARGS;
  if (!ok(env, dio_prefix_sum_encode(handle(&c, 0)))) return NULL;
  return undefined(env);
}
FN(prefix_sum_destroy) {
  ARGS;
  dio_prefix_sum_destroy(handle(&c, 0));
  return undefined(env);
}

/* ----------------------------------------------------------- RadixSort -- */

FN(radix_sort_create) {
  ARGS;
  dio_radix_sort* s = NULL;
  if (!ok(env, dio_radix_sort_create(handle(&c, 0), u32(&c, 1), u32(&c, 2), handle(&c, 3), &s))) return NULL;
  return wrap(env, s);
}
FN(radix_sort_keys) {
  ARGS;
  return wrap(env, dio_radix_sort_keys(handle(&c, 0)));
}
FN(radix_sort_values) {
  ARGS;
  return wrap(env, dio_radix_sort_values(handle(&c, 0)));
}
FN(radix_sort_capacity) {
  ARGS;
  return number(env, dio_radix_sort_capacity(handle(&c, 0)));
}
FN(radix_sort_encode) {
  ARGS;
  if (!ok(env, dio_radix_sort_encode(handle(&c, 0)))) return NULL;
  return undefined(env);
}
FN(radix_sort_destroy) {
  ARGS;
  dio_radix_sort_destroy(handle(&c, 0));
  return undefined(env);
}

/* ------------------------------------------------------ TileRasterizer -- */

static int background_of(call_t* c, size_t i, float bg[3]) { return floats(c, c->argv[i], bg, 3); }

FN(tiles_create) {
  ARGS;
  dio_tile_rasterizer* t = NULL;
  if (!ok(env, dio_tile_rasterizer_create(handle(&c, 0), handle(&c, 1), u32(&c, 2), u32(&c, 3), &t))) return NULL;
  return wrap(env, t);
}
FN(tiles_encode) {
  ARGS;
  float bg[3];
  if (!background_of(&c, 4, bg)) return fail(env, "background must be [r, g, b]");
  if (!ok(env, dio_tile_rasterizer_encode(handle(&c, 0), handle(&c, 1), u32(&c, 2), u32(&c, 3), bg))) return NULL;
  return undefined(env);
}
FN(tiles_encode_backward) {
  ARGS;
  if (!ok(env, dio_tile_rasterizer_encode_backward(handle(&c, 0), u32(&c, 1), u32(&c, 2), handle(&c, 3),
                                                   handle(&c, 4))))
    return NULL;
  return undefined(env);
}
FN(tiles_after_submit) {
  ARGS;
  if (!ok(env, dio_tile_rasterizer_after_submit(handle(&c, 0)))) return NULL;
  return undefined(env);
}
FN(tiles_last_pair_count) {
  ARGS;
  return number(env, dio_tile_rasterizer_last_pair_count(handle(&c, 0)));
}
FN(tiles_pair_capacity) {
  ARGS;
  return number(env, dio_tile_rasterizer_pair_capacity(handle(&c, 0)));
}
FN(tiles_pair_splats) {
  ARGS;
  return wrap(env, dio_tile_rasterizer_pair_splats(handle(&c, 0)));
}
FN(tiles_ranges) {
  ARGS;
  return wrap(env, dio_tile_rasterizer_ranges(handle(&c, 0)));
}
FN(tiles_pixel_color) {
  ARGS;
  return wrap(env, dio_tile_rasterizer_pixel_color(handle(&c, 0)));
}
FN(tiles_pixel_last) {
  ARGS;
  return wrap(env, dio_tile_rasterizer_pixel_last(handle(&c, 0)));
}
FN(tiles_destroy) {
  ARGS;
  dio_tile_rasterizer_destroy(handle(&c, 0));
  return undefined(env);
}

/* ---------------------------------------------------- GaussianRenderer -- */

FN(renderer_create) {
  ARGS;
  dio_renderer* r = NULL;
  if (!ok(env, dio_renderer_create(handle(&c, 0), u32(&c, 1), &r))) return NULL;
  return wrap(env, r);
}
FN(renderer_set_gaussians) {
  ARGS;
  void *g = NULL, *sh = NULL;
  size_t glen = 0, shlen = 0;
  if (!bytes_of(&c, 1, &g, &glen)) return fail(env, "setGaussians needs a Float32Array");
  if (!is_nullish(&c, 2) && !bytes_of(&c, 2, &sh, &shlen)) return fail(env, "sh rest must be a Float32Array");
  uint32_t count = (uint32_t)(glen / 64);
  if (sh && shlen != (size_t)count * 45 * 4) {
    char message[160];
    snprintf(message, sizeof message, "Expected %u spherical-harmonic floats, got %zu.", count * 45, shlen / 4);
    return fail(env, message);
  }
  if (!ok(env, dio_renderer_set_gaussians(handle(&c, 0), (const float*)g, count, (const float*)sh, u32(&c, 3))))
    return NULL;
  return undefined(env);
}
FN(renderer_adopt_gaussians) {
  ARGS;
  if (!ok(env, dio_renderer_adopt_gaussians(handle(&c, 0), u32(&c, 1), handle(&c, 2), handle(&c, 3), u32(&c, 4))))
    return NULL;
  return undefined(env);
}
FN(renderer_set_sh_degree) {
  ARGS;
  if (!ok(env, dio_renderer_set_sh_degree(handle(&c, 0), u32(&c, 1)))) return NULL;
  return undefined(env);
}
FN(renderer_encode_preprocess) {
  ARGS;
  dio_camera cam;
  if (!camera_of(&c, c.argv[1], &cam)) return fail(env, "camera must be {view, proj, viewport}");
  if (!ok(env, dio_renderer_encode_preprocess(handle(&c, 0), &cam))) return NULL;
  return undefined(env);
}
FN(renderer_encode_tiles) {
  ARGS;
  float bg[3];
  if (!background_of(&c, 2, bg)) return fail(env, "background must be [r, g, b]");
  if (!ok(env, dio_renderer_encode_tiles(handle(&c, 0), handle(&c, 1), bg))) return NULL;
  return undefined(env);
}
FN(renderer_encode_draw) {
  ARGS;
  float bg[3];
  if (!background_of(&c, 6, bg)) return fail(env, "background must be [r, g, b]");
  if (!ok(env, dio_renderer_encode_draw(handle(&c, 0), handle(&c, 1), handle(&c, 2), u32(&c, 3), u32(&c, 4),
                                        u32(&c, 5), bg)))
    return NULL;
  return undefined(env);
}
FN(renderer_encode_backward) {
  ARGS;
  if (!ok(env, dio_renderer_encode_backward(handle(&c, 0), handle(&c, 1)))) return NULL;
  return undefined(env);
}
FN(renderer_after_submit) {
  ARGS;
  if (!ok(env, dio_renderer_after_submit(handle(&c, 0)))) return NULL;
  return undefined(env);
}
FN(renderer_count) {
  ARGS;
  return number(env, dio_renderer_count(handle(&c, 0)));
}
#define GETTER(fn, call) \
  FN(fn) {               \
    ARGS;                \
    return wrap(env, call(handle(&c, 0))); \
  }
GETTER(renderer_splat_buffer, dio_renderer_splat_buffer)
GETTER(renderer_gaussian_buffer, dio_renderer_gaussian_buffer)
GETTER(renderer_sh_buffer, dio_renderer_sh_buffer)
GETTER(renderer_tiles, dio_renderer_tiles)
GETTER(renderer_pixel_buffer, dio_renderer_pixel_buffer)
GETTER(renderer_pixel_last_buffer, dio_renderer_pixel_last_buffer)
GETTER(renderer_gradient_gaussians, dio_renderer_gradient_gaussians)
GETTER(renderer_gradient_sh, dio_renderer_gradient_sh)
GETTER(renderer_gradient_splats, dio_renderer_gradient_splats)
FN(renderer_destroy) {
  ARGS;
  dio_renderer_destroy(handle(&c, 0));
  return undefined(env);
}

/* -------------------------------------------------------- TrainingLoss -- */

FN(loss_create) {
  ARGS;
  dio_training_loss* l = NULL;
  if (!ok(env, dio_training_loss_create(handle(&c, 0), u32(&c, 1), u32(&c, 2), &l))) return NULL;
  return wrap(env, l);
}
FN(loss_encode) {
  ARGS;
  if (!ok(env, dio_training_loss_encode(handle(&c, 0), handle(&c, 1), handle(&c, 2)))) return NULL;
  return undefined(env);
}
FN(loss_read) {
  ARGS;
  double v = 0;
  if (!ok(env, dio_training_loss_read(handle(&c, 0), &v))) return NULL;
  return number(env, v);
}
GETTER(loss_pixel_grads, dio_training_loss_pixel_grads)
GETTER(loss_pixel_loss, dio_training_loss_pixel_loss)
FN(loss_destroy) {
  ARGS;
  dio_training_loss_destroy(handle(&c, 0));
  return undefined(env);
}

/* ------------------------------------------------------------- Trainer -- */

/* Field names of TrainingSettings in src/trainer.ts:37-63 order. */
static napi_value settings_to_js(napi_env env, const dio_training_settings* s) {
  napi_value o, arr;
  napi_create_object(env, &o);
#define SET(name, value) napi_set_named_property(env, o, name, number(env, value))
  SET("extent", s->extent);
  SET("positionLrSteps", s->position_lr_steps);
  napi_create_array_with_length(env, 2, &arr);
  napi_set_element(env, arr, 0, number(env, s->position_lr[0]));
  napi_set_element(env, arr, 1, number(env, s->position_lr[1]));
  napi_set_named_property(env, o, "positionLr", arr);
  SET("opacityLr", s->opacity_lr);
  SET("scaleLr", s->scale_lr);
  SET("rotationLr", s->rotation_lr);
  SET("colorLr", s->color_lr);
  SET("shLr", s->sh_lr);
  SET("shDegreeInterval", s->sh_degree_interval);
  SET("densifyFrom", s->densify_from);
  SET("densifyUntil", s->densify_until);
  SET("densifyInterval", s->densify_interval);
  SET("densifyGradThreshold", s->densify_grad_threshold);
  SET("percentDense", s->percent_dense);
  SET("minOpacity", s->min_opacity);
  SET("opacityResetInterval", s->opacity_reset_interval);
#undef SET
  napi_create_array_with_length(env, 3, &arr);
  for (uint32_t k = 0; k < 3; k++) napi_set_element(env, arr, k, number(env, s->background[k]));
  napi_set_named_property(env, o, "background", arr);
  return o;
}

static double num_prop(napi_env env, napi_value o, const char* name) {
  double d = 0;
  napi_get_value_double(env, property(env, o, name), &d);
  return d;
}

static int settings_from_js(call_t* c, napi_value o, dio_training_settings* s) {
  napi_env env = c->env;
  s->extent = (float)num_prop(env, o, "extent");
  s->position_lr_steps = (uint32_t)num_prop(env, o, "positionLrSteps");
  if (!floats(c, property(env, o, "positionLr"), s->position_lr, 2)) return 0;
  s->opacity_lr = (float)num_prop(env, o, "opacityLr");
  s->scale_lr = (float)num_prop(env, o, "scaleLr");
  s->rotation_lr = (float)num_prop(env, o, "rotationLr");
  s->color_lr = (float)num_prop(env, o, "colorLr");
  s->sh_lr = (float)num_prop(env, o, "shLr");
  s->sh_degree_interval = (uint32_t)num_prop(env, o, "shDegreeInterval");
  s->densify_from = (uint32_t)num_prop(env, o, "densifyFrom");
  s->densify_until = (uint32_t)num_prop(env, o, "densifyUntil");
  s->densify_interval = (uint32_t)num_prop(env, o, "densifyInterval");
  s->densify_grad_threshold = (float)num_prop(env, o, "densifyGradThreshold");
  s->percent_dense = (float)num_prop(env, o, "percentDense");
  s->min_opacity = (float)num_prop(env, o, "minOpacity");
  s->opacity_reset_interval = (uint32_t)num_prop(env, o, "opacityResetInterval");
  return floats(c, property(env, o, "background"), s->background, 3);
}

FN(default_settings) {
  ARGS;
  dio_training_settings s;
  dio_default_settings((float)f64(&c, 0), &s);
  return settings_to_js(env, &s);
}

/* Views: [{camera, image: Uint32Array rgba8}] (src/trainer.ts:20-24). */
static int views_of(call_t* c, napi_value arr, dio_camera** cams, const uint32_t*** images, uint32_t* count) {
  napi_get_array_length(c->env, arr, count);
  *cams = (dio_camera*)calloc(*count ? *count : 1, sizeof(dio_camera));
  *images = (const uint32_t**)calloc(*count ? *count : 1, sizeof(uint32_t*));
  for (uint32_t k = 0; k < *count; k++) {
    napi_value v, img;
    napi_get_element(c->env, arr, k, &v);
    if (!camera_of(c, property(c->env, v, "camera"), &(*cams)[k])) return 0;
    img = property(c->env, v, "image");
    call_t sub = *c;
    sub.argc = 1;
    sub.argv[0] = img;
    void* data;
    size_t length;
    if (!bytes_of(&sub, 0, &data, &length)) return 0;
    (*images)[k] = (const uint32_t*)data;
  }
  return 1;
}

FN(trainer_create) {
  ARGS;
  void *params, *sh;
  size_t plen, shlen;
  if (!bytes_of(&c, 1, &params, &plen) || !bytes_of(&c, 2, &sh, &shlen))
    return fail(env, "raw params and sh must be Float32Arrays");
  dio_camera* cams = NULL;
  const uint32_t** images = NULL;
  uint32_t views = 0;
  dio_training_settings s;
  napi_value result = NULL;
  if (!views_of(&c, c.argv[3], &cams, &images, &views)) {
    fail(env, "views must be [{camera, image}]");
  } else if (!settings_from_js(&c, c.argv[4], &s)) {
    fail(env, "settings are incomplete");
  } else {
    dio_trainer* t = NULL;
    const uint32_t batch = u32(&c, 5); /* optional: views per step of a batch trainer */
    const int status = batch ? dio_trainer_create_batched(handle(&c, 0), (const float*)params, (const float*)sh,
                                                          (uint32_t)(plen / 64), cams, images, views, batch, &s, &t)
                             : dio_trainer_create(handle(&c, 0), (const float*)params, (const float*)sh,
                                                  (uint32_t)(plen / 64), cams, images, views, &s, &t);
    if (ok(env, status)) result = wrap(env, t);
  }
  free(cams);
  free((void*)images);
  return result;
}
#define TRAINER_VOID(fn, call) \
  FN(fn) {                     \
    ARGS;                      \
    if (!ok(env, call(handle(&c, 0)))) return NULL; \
    return undefined(env);     \
  }
TRAINER_VOID(trainer_step, dio_trainer_step)
TRAINER_VOID(trainer_optimize, dio_trainer_optimize)
TRAINER_VOID(trainer_maintain, dio_trainer_maintain)
TRAINER_VOID(trainer_reset_opacity, dio_trainer_reset_opacity)
FN(trainer_densify) {
  ARGS;
  bool prune = false;
  napi_get_value_bool(env, c.argv[1], &prune);
  if (!ok(env, dio_trainer_densify(handle(&c, 0), prune ? 1 : 0))) return NULL;
  return undefined(env);
}
FN(trainer_export_raw) {
  ARGS;
  dio_trainer* t = handle(&c, 0);
  uint32_t n = dio_trainer_count(t);
  void *p, *s;
  napi_value pab, sab, pta, sta, o;
  napi_create_arraybuffer(env, (size_t)n * 64, &p, &pab);
  napi_create_arraybuffer(env, (size_t)n * 180, &s, &sab);
  if (!ok(env, dio_trainer_export_raw(t, (float*)p, (float*)s))) return NULL;
  napi_create_typedarray(env, napi_float32_array, (size_t)n * 16, pab, 0, &pta);
  napi_create_typedarray(env, napi_float32_array, (size_t)n * 45, sab, 0, &sta);
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "params", pta);
  napi_set_named_property(env, o, "sh", sta);
  return o;
}
FN(trainer_last_loss) {
  ARGS;
  double v = 0;
  if (!ok(env, dio_trainer_last_loss(handle(&c, 0), &v))) return NULL;
  return number(env, v);
}
FN(trainer_evaluate) {
  ARGS;
  dio_camera* cams = NULL;
  const uint32_t** images = NULL;
  uint32_t views = 0;
  bool with_ssim = true;
  napi_value result = NULL;
  if (!is_nullish(&c, 2)) napi_get_value_bool(env, c.argv[2], &with_ssim);
  if (!views_of(&c, c.argv[1], &cams, &images, &views)) {
    fail(env, "views must be [{camera, image}]");
  } else {
    double psnr = 0, ssim = 0;
    if (ok(env, dio_trainer_evaluate(handle(&c, 0), cams, images, views, with_ssim ? 1 : 0, &psnr, &ssim))) {
      napi_create_object(env, &result);
      napi_set_named_property(env, result, "psnr", number(env, psnr));
      napi_set_named_property(env, result, "ssim", number(env, ssim));
    }
  }
  free(cams);
  free((void*)images);
  return result;
}
/* measurePsnr: (trainer, [camera], [image buffer handle]) -> mean PSNR */
FN(trainer_measure_psnr) {
  ARGS;
  uint32_t views = 0, handles = 0;
  napi_get_array_length(env, c.argv[1], &views);
  napi_get_array_length(env, c.argv[2], &handles);
  if (views != handles) return fail(env, "one image per camera");
  dio_camera* cams = (dio_camera*)calloc(views ? views : 1, sizeof(dio_camera));
  const dio_buf** images = (const dio_buf**)calloc(views ? views : 1, sizeof(dio_buf*));
  napi_value result = NULL;
  int good = 1;
  for (uint32_t k = 0; k < views && good; k++) {
    napi_value v;
    napi_get_element(env, c.argv[1], k, &v);
    good = camera_of(&c, v, &cams[k]);
    napi_get_element(env, c.argv[2], k, &v);
    void* h = NULL;
    if (good) good = napi_get_value_external(env, v, &h) == napi_ok;
    images[k] = (const dio_buf*)h;
  }
  double psnr = 0;
  if (!good) {
    fail(env, "views must be [{camera, image: Buffer}]");
  } else if (ok(env, dio_trainer_measure_psnr(handle(&c, 0), cams, images, views, &psnr))) {
    result = number(env, psnr);
  }
  free(cams);
  free((void*)images);
  return result;
}
FN(trainer_count) {
  ARGS;
  return number(env, dio_trainer_count(handle(&c, 0)));
}
FN(trainer_iteration) {
  ARGS;
  return number(env, dio_trainer_iteration(handle(&c, 0)));
}
FN(trainer_set_iteration) {
  ARGS;
  dio_trainer_set_iteration(handle(&c, 0), u32(&c, 1));
  return undefined(env);
}
FN(trainer_sh_degree) {
  ARGS;
  return number(env, dio_trainer_sh_degree(handle(&c, 0)));
}
FN(trainer_set_sh_degree) {
  ARGS;
  if (!ok(env, dio_trainer_set_sh_degree(handle(&c, 0), u32(&c, 1)))) return NULL;
  return undefined(env);
}
FN(trainer_densifications) {
  ARGS;
  dio_trainer* t = handle(&c, 0);
  uint32_t n = dio_trainer_densification_count(t);
  napi_value arr;
  napi_create_array_with_length(env, n, &arr);
  for (uint32_t k = 0; k < n; k++) {
    dio_densification d;
    napi_value o;
    if (!ok(env, dio_trainer_densification(t, k, &d))) return NULL;
    napi_create_object(env, &o);
    napi_set_named_property(env, o, "step", number(env, d.step));
    napi_set_named_property(env, o, "before", number(env, d.before));
    napi_set_named_property(env, o, "after", number(env, d.after));
    napi_set_named_property(env, o, "kept", number(env, d.kept));
    napi_set_named_property(env, o, "cloned", number(env, d.cloned));
    napi_set_named_property(env, o, "split", number(env, d.split));
    napi_set_named_property(env, o, "removed", number(env, d.removed));
    napi_set_element(env, arr, k, o);
  }
  return arr;
}
FN(trainer_batch_size) {
  ARGS;
  return number(env, dio_trainer_batch_size(handle(&c, 0)));
}
GETTER(trainer_params, dio_trainer_params)
GETTER(trainer_moments, dio_trainer_moments)
GETTER(trainer_stats, dio_trainer_stats)
GETTER(trainer_renderer, dio_trainer_renderer)
FN(trainer_destroy) {
  ARGS;
  dio_trainer_destroy(handle(&c, 0));
  return undefined(env);
}


/* ------------------------------------------------ learned densification -- */

static napi_value float_array(napi_env env, const float* src, size_t n) {
  void* data;
  napi_value ab, ta;
  napi_create_arraybuffer(env, n * 4, &data, &ab);
  memcpy(data, src, n * 4);
  napi_create_typedarray(env, napi_float32_array, n, ab, 0, &ta);
  return ta;
}

FN(mlp_create) {
  ARGS;
  dio_densify_mlp* m = NULL;
  if (!ok(env, dio_mlp_create(handle(&c, 0), u32(&c, 1), &m))) return NULL;
  return wrap(env, m);
}
FN(mlp_get_weights) {
  ARGS;
  float w[DIO_MLP_PARAMS];
  if (!ok(env, dio_mlp_get_weights(handle(&c, 0), w))) return NULL;
  return float_array(env, w, DIO_MLP_PARAMS);
}
FN(mlp_set_weights) {
  ARGS;
  void* w = NULL;
  size_t len = 0;
  if (!bytes_of(&c, 1, &w, &len) || len != DIO_MLP_PARAMS * 4) return fail(env, "weights must be 5252 floats");
  if (!ok(env, dio_mlp_set_weights(handle(&c, 0), (const float*)w))) return NULL;
  return undefined(env);
}
FN(mlp_set_input_norm) {
  ARGS;
  void *mean = NULL, *scale = NULL;
  size_t ml = 0, sl = 0;
  if (!bytes_of(&c, 1, &mean, &ml) || !bytes_of(&c, 2, &scale, &sl) || ml != 48 || sl != 48)
    return fail(env, "mean and scale must be 12 floats each");
  if (!ok(env, dio_mlp_set_input_norm(handle(&c, 0), (const float*)mean, (const float*)scale))) return NULL;
  return undefined(env);
}
FN(mlp_set_class_weights) {
  ARGS;
  float w[4];
  if (!floats(&c, c.argv[1], w, 4)) return fail(env, "class weights must be 4 numbers");
  if (!ok(env, dio_mlp_set_class_weights(handle(&c, 0), w))) return NULL;
  return undefined(env);
}
FN(mlp_forward) {
  ARGS;
  if (!ok(env, dio_mlp_forward(handle(&c, 0), handle(&c, 1), u32(&c, 2), handle(&c, 3)))) return NULL;
  return undefined(env);
}
FN(mlp_gradients) {
  ARGS;
  float g[DIO_MLP_PARAMS];
  double loss = 0;
  napi_value o;
  if (!ok(env, dio_mlp_gradients(handle(&c, 0), handle(&c, 1), handle(&c, 2), u32(&c, 3), g, &loss))) return NULL;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "grads", float_array(env, g, DIO_MLP_PARAMS));
  napi_set_named_property(env, o, "loss", number(env, loss));
  return o;
}
FN(mlp_train) {
  ARGS;
  double loss = 0;
  if (!ok(env, dio_mlp_train(handle(&c, 0), handle(&c, 1), handle(&c, 2), u32(&c, 3), u32(&c, 4), u32(&c, 5),
                             (float)f64(&c, 6), u32(&c, 7), &loss)))
    return NULL;
  return number(env, loss);
}
FN(mlp_destroy) {
  ARGS;
  dio_mlp_destroy(handle(&c, 0));
  return undefined(env);
}
FN(trainer_set_densify_mode) {
  ARGS;
  if (!ok(env, dio_trainer_set_densify_mode(handle(&c, 0), u32(&c, 1), handle(&c, 2), u32(&c, 3), (float)f64(&c, 4),
                                            u32(&c, 5))))
    return NULL;
  return undefined(env);
}
FN(trainer_samples) {
  ARGS;
  dio_buf *f = NULL, *l = NULL;
  uint32_t n = 0;
  napi_value o;
  if (!ok(env, dio_trainer_samples(handle(&c, 0), &f, &l, &n))) return NULL;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "features", wrap(env, f));
  napi_set_named_property(env, o, "labels", wrap(env, l));
  napi_set_named_property(env, o, "count", number(env, n));
  return o;
}
FN(trainer_features) {
  ARGS;
  dio_buf* f = NULL;
  if (!ok(env, dio_trainer_features(handle(&c, 0), &f))) return NULL;
  return wrap(env, f);
}
GETTER(trainer_extra_stats, dio_trainer_extra_stats)

FN(mean_neighbor_distance2) {
  ARGS;
  if (!ok(env, dio_mean_neighbor_distance2(handle(&c, 0), handle(&c, 1), u32(&c, 2), handle(&c, 3)))) return NULL;
  return undefined(env);
}
FN(initial_gaussians) {
  ARGS;
  if (!ok(env, dio_initial_gaussians(handle(&c, 0), handle(&c, 1), handle(&c, 2), u32(&c, 3), handle(&c, 4),
                                     handle(&c, 5))))
    return NULL;
  return undefined(env);
}
/* ------------------------------------------------------------- COLMAP */

static napi_value double_list(napi_env env, const double* v, uint32_t n) {
  napi_value arr;
  napi_create_array_with_length(env, n, &arr);
  for (uint32_t k = 0; k < n; k++) napi_set_element(env, arr, k, number(env, v[k]));
  return arr;
}

static int doubles(call_t* c, napi_value arr, double* out, uint32_t n) {
  for (uint32_t k = 0; k < n; k++) {
    napi_value e;
    if (napi_get_element(c->env, arr, k, &e) != napi_ok) return 0;
    if (napi_get_value_double(c->env, e, &out[k]) != napi_ok) return 0;
  }
  return 1;
}

/* parseColmap: (camerasBin, imagesBin, pointsBin) -> {cameras, positions, colors} */
FN(colmap_parse) {
  ARGS;
  void *cb = NULL, *ib = NULL, *pb = NULL;
  size_t cn = 0, in = 0, pn = 0;
  if (!bytes_of(&c, 0, &cb, &cn) || !bytes_of(&c, 1, &ib, &in) || !bytes_of(&c, 2, &pb, &pn))
    return fail(env, "parseColmap needs three ArrayBuffers");
  uint32_t ncam = 0, npts = 0;
  if (!ok(env, dio_colmap_parse(cb, cn, ib, in, pb, pn, NULL, 0, &ncam, NULL, NULL, 0, &npts))) return NULL;
  dio_colmap_camera* cams = (dio_colmap_camera*)calloc(ncam ? ncam : 1, sizeof(dio_colmap_camera));
  void *pos_data = NULL, *col_data = NULL;
  napi_value pos_ab, col_ab, positions, colors, result = NULL;
  napi_create_arraybuffer(env, (size_t)npts * 12, &pos_data, &pos_ab);
  napi_create_arraybuffer(env, (size_t)npts * 3, &col_data, &col_ab);
  if (ok(env, dio_colmap_parse(cb, cn, ib, in, pb, pn, cams, ncam, &ncam, (float*)pos_data, (uint8_t*)col_data, npts,
                               &npts))) {
    napi_create_typedarray(env, napi_float32_array, (size_t)npts * 3, pos_ab, 0, &positions);
    napi_create_typedarray(env, napi_uint8_array, (size_t)npts * 3, col_ab, 0, &colors);
    napi_value list;
    napi_create_array_with_length(env, ncam, &list);
    for (uint32_t k = 0; k < ncam; k++) {
      napi_value o, name;
      napi_create_object(env, &o);
      napi_create_string_utf8(env, cams[k].name, NAPI_AUTO_LENGTH, &name);
      napi_set_named_property(env, o, "name", name);
      napi_set_named_property(env, o, "width", number(env, cams[k].width));
      napi_set_named_property(env, o, "height", number(env, cams[k].height));
      napi_set_named_property(env, o, "position", double_list(env, cams[k].position, 3));
      napi_set_named_property(env, o, "right", double_list(env, cams[k].right, 3));
      napi_set_named_property(env, o, "down", double_list(env, cams[k].down, 3));
      napi_set_named_property(env, o, "forward", double_list(env, cams[k].forward, 3));
      napi_set_named_property(env, o, "fx", number(env, cams[k].fx));
      napi_set_named_property(env, o, "fy", number(env, cams[k].fy));
      napi_set_element(env, list, k, o);
    }
    napi_create_object(env, &result);
    napi_set_named_property(env, result, "cameras", list);
    napi_set_named_property(env, result, "positions", positions);
    napi_set_named_property(env, result, "colors", colors);
  }
  free(cams);
  return result;
}

/* trainingCameraData: (TrainingCamera, width, height) -> {view, proj, viewport} */
FN(colmap_training_camera) {
  ARGS;
  dio_colmap_camera cam;
  memset(&cam, 0, sizeof cam);
  napi_value o = c.argv[0];
  double w = 0, h = 0;
  if (!doubles(&c, property(env, o, "position"), cam.position, 3) ||
      !doubles(&c, property(env, o, "right"), cam.right, 3) || !doubles(&c, property(env, o, "down"), cam.down, 3) ||
      !doubles(&c, property(env, o, "forward"), cam.forward, 3))
    return fail(env, "a TrainingCamera needs position, right, down and forward");
  napi_get_value_double(env, property(env, o, "fx"), &cam.fx);
  napi_get_value_double(env, property(env, o, "fy"), &cam.fy);
  napi_get_value_double(env, property(env, o, "width"), &w);
  napi_get_value_double(env, property(env, o, "height"), &h);
  cam.width = (uint32_t)w;
  cam.height = (uint32_t)h;
  dio_camera out;
  if (!ok(env, dio_colmap_training_camera(&cam, u32(&c, 1), u32(&c, 2), &out))) return NULL;
  napi_value r, view, proj, viewport;
  napi_create_object(env, &r);
  napi_create_array_with_length(env, 16, &view);
  napi_create_array_with_length(env, 16, &proj);
  napi_create_array_with_length(env, 2, &viewport);
  for (uint32_t k = 0; k < 16; k++) {
    napi_set_element(env, view, k, number(env, out.view[k]));
    napi_set_element(env, proj, k, number(env, out.proj[k]));
  }
  napi_set_element(env, viewport, 0, number(env, out.viewport[0]));
  napi_set_element(env, viewport, 1, number(env, out.viewport[1]));
  napi_set_named_property(env, r, "view", view);
  napi_set_named_property(env, r, "proj", proj);
  napi_set_named_property(env, r, "viewport", viewport);
  return r;
}

/* writeGaussianPly: (params, sh, count) -> ArrayBuffer holding the file. */
FN(ply_write) {
  ARGS;
  size_t length = 0;
  void* data = NULL;
  napi_value ab;
  if (!ok(env, dio_ply_write(NULL, NULL, u32(&c, 2), NULL, 0, &length))) return NULL;
  if (napi_create_arraybuffer(env, length, &data, &ab) != napi_ok) return fail(env, "could not allocate the .ply");
  if (!ok(env, dio_ply_write(handle(&c, 0), handle(&c, 1), u32(&c, 2), data, length, &length))) return NULL;
  return ab;
}
/* parseGaussianPly: (ctx, ArrayBuffer | TypedArray) -> {gaussians, shRest, count, shDegree} with owned buffers. */
FN(ply_parse) {
  ARGS;
  void* data = NULL;
  size_t length = 0;
  dio_buf *g = NULL, *s = NULL;
  uint32_t n = 0, d = 0;
  napi_value o;
  if (!bytes_of(&c, 1, &data, &length)) return fail(env, "parseGaussianPly needs an ArrayBuffer or TypedArray");
  if (!ok(env, dio_ply_parse(handle(&c, 0), data, length, &g, &s, &n, &d))) return NULL;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "gaussians", wrap(env, g));
  napi_set_named_property(env, o, "shRest", wrap(env, s));
  napi_set_named_property(env, o, "count", number(env, n));
  napi_set_named_property(env, o, "shDegree", number(env, d));
  return o;
}


/* --------------------------------------------------------------- tuning -- */

static napi_value tuning_object(napi_env env, const dio_tuning* k) {
  napi_value o;
  napi_create_object(env, &o);
  napi_set_named_property(env, o, "tileSize", number(env, k->tile_size));
  napi_set_named_property(env, o, "splatBatch", number(env, k->splat_batch));
  napi_set_named_property(env, o, "digitBits", number(env, k->digit_bits));
  napi_set_named_property(env, o, "persistent", number(env, k->persistent));
  napi_set_named_property(env, o, "segment", number(env, k->segment));
  napi_set_named_property(env, o, "perTileDepth", number(env, k->per_tile_depth));
  return o;
}

FN(trainer_get_tuning) {
  ARGS;
  dio_tuning k;
  if (!ok(env, dio_trainer_get_tuning(handle(&c, 0), &k))) return NULL;
  return tuning_object(env, &k);
}
static int read_tuning(call_t* c, napi_value arr, dio_tuning* k) {
  float v[6];
  if (!floats(c, arr, v, 6)) return 0;
  k->tile_size = (uint32_t)v[0];
  k->splat_batch = (uint32_t)v[1];
  k->digit_bits = (uint32_t)v[2];
  k->persistent = (uint32_t)v[3];
  k->segment = (uint32_t)v[4];
  k->per_tile_depth = (uint32_t)v[5];
  return 1;
}
FN(trainer_set_tuning) {
  ARGS;
  dio_tuning k;
  if (!read_tuning(&c, c.argv[1], &k))
    return fail(env, "tuning must be [tileSize, splatBatch, digitBits, persistent, segment, perTileDepth]");
  if (!ok(env, dio_trainer_set_tuning(handle(&c, 0), &k))) return NULL;
  return undefined(env);
}
FN(tiles_set_tuning) {
  ARGS;
  dio_tuning k;
  if (!read_tuning(&c, c.argv[1], &k))
    return fail(env, "tuning must be [tileSize, splatBatch, digitBits, persistent, segment, perTileDepth]");
  if (!ok(env, dio_tile_rasterizer_set_tuning(handle(&c, 0), &k))) return NULL;
  return undefined(env);
}
FN(trainer_autotune) {
  ARGS;
  dio_tuning_result r[128];
  uint32_t n = 0;
  napi_value arr;
  if (!ok(env, dio_trainer_autotune(handle(&c, 0), u32(&c, 1), r, 128, &n))) return NULL;
  if (n > 128) n = 128;
  napi_create_array_with_length(env, n, &arr);
  for (uint32_t k = 0; k < n; k++) {
    napi_value o = tuning_object(env, &r[k].tuning);
    napi_set_named_property(env, o, "msPerStep", number(env, r[k].ms_per_step));
    napi_set_element(env, arr, k, o);
  }
  return arr;
}

/* ---------------------------------------------------------------- init -- */

#define EXPORT(name) {#name, NULL, name, NULL, NULL, NULL, napi_default, NULL}

static napi_value init(napi_env env, napi_value exports) {
  napi_property_descriptor d[] = {
      EXPORT(version), EXPORT(ctx_create), EXPORT(ctx_synchronize), EXPORT(ctx_destroy), EXPORT(buf_create),
      EXPORT(buf_upload), EXPORT(buf_download), EXPORT(buf_clear), EXPORT(buf_copy), EXPORT(ctx_timestamp), EXPORT(ctx_elapsed), EXPORT(buf_size), EXPORT(buf_destroy),
      EXPORT(prefix_sum_create), EXPORT(prefix_sum_total), EXPORT(prefix_sum_encode), EXPORT(prefix_sum_destroy),
      EXPORT(radix_sort_create), EXPORT(radix_sort_keys), EXPORT(radix_sort_values), EXPORT(radix_sort_capacity),
      EXPORT(radix_sort_encode), EXPORT(radix_sort_destroy), EXPORT(tiles_create), EXPORT(tiles_encode),
      EXPORT(tiles_encode_backward), EXPORT(tiles_after_submit), EXPORT(tiles_set_tuning), EXPORT(tiles_last_pair_count),
      EXPORT(tiles_pair_capacity), EXPORT(tiles_pair_splats), EXPORT(tiles_ranges), EXPORT(tiles_pixel_color),
      EXPORT(tiles_pixel_last), EXPORT(tiles_destroy), EXPORT(renderer_create), EXPORT(renderer_set_gaussians),
      EXPORT(renderer_adopt_gaussians), EXPORT(renderer_set_sh_degree), EXPORT(renderer_encode_preprocess),
      EXPORT(renderer_encode_tiles), EXPORT(renderer_encode_draw), EXPORT(renderer_encode_backward),
      EXPORT(renderer_after_submit), EXPORT(renderer_count), EXPORT(renderer_splat_buffer),
      EXPORT(renderer_gaussian_buffer), EXPORT(renderer_sh_buffer), EXPORT(renderer_tiles),
      EXPORT(renderer_pixel_buffer), EXPORT(renderer_pixel_last_buffer), EXPORT(renderer_gradient_gaussians),
      EXPORT(renderer_gradient_sh), EXPORT(renderer_gradient_splats), EXPORT(renderer_destroy), EXPORT(loss_create),
      EXPORT(loss_encode), EXPORT(loss_read), EXPORT(loss_pixel_grads), EXPORT(loss_pixel_loss), EXPORT(loss_destroy),
      EXPORT(default_settings), EXPORT(trainer_create), EXPORT(trainer_step), EXPORT(trainer_optimize),
      EXPORT(trainer_maintain), EXPORT(trainer_reset_opacity), EXPORT(trainer_densify), EXPORT(trainer_export_raw),
      EXPORT(trainer_last_loss), EXPORT(trainer_evaluate), EXPORT(trainer_count), EXPORT(trainer_iteration),
      EXPORT(trainer_set_iteration), EXPORT(trainer_sh_degree), EXPORT(trainer_set_sh_degree),
      EXPORT(trainer_densifications), EXPORT(trainer_params), EXPORT(trainer_moments), EXPORT(trainer_stats),
      EXPORT(trainer_renderer), EXPORT(trainer_destroy), EXPORT(trainer_batch_size),
      EXPORT(mlp_create), EXPORT(mlp_get_weights), EXPORT(mlp_set_weights), EXPORT(mlp_set_input_norm),
      EXPORT(mlp_set_class_weights), EXPORT(mlp_forward), EXPORT(mlp_gradients), EXPORT(mlp_train),
      EXPORT(mlp_destroy), EXPORT(trainer_set_densify_mode), EXPORT(trainer_samples), EXPORT(trainer_features),
      EXPORT(trainer_extra_stats), EXPORT(trainer_get_tuning), EXPORT(trainer_set_tuning), EXPORT(trainer_autotune),
      EXPORT(ply_write), EXPORT(ply_parse), EXPORT(mean_neighbor_distance2), EXPORT(initial_gaussians),
      EXPORT(trainer_measure_psnr), EXPORT(colmap_parse), EXPORT(colmap_training_camera),
  };
  napi_define_properties(env, exports, sizeof d / sizeof d[0], d);
  return exports;
}

NAPI_MODULE(NODE_GYP_MODULE_NAME, init)
