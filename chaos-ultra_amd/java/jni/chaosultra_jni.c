/*
 * chaosultra_jni.c — JNI glue between hiprenderer.HipNative and the C ABI of
 * libchaosultra_hip (include/chaosultra.h).  One function per native method;
 * a non-CU_OK status throws the Java exception class the status stands for,
 * with cu_last_error() as its message.
 *
 * Build (wherever a JDK is installed; JAVA_HOME points at it):
 *   cc -O2 -fPIC -shared chaosultra_jni.c -o libchaosultra_jni.so \
 *      -I"$JAVA_HOME/include" -I"$JAVA_HOME/include/linux" -I../../../include \
 *      -L../../lib -lchaosultra_hip -Wl,-rpath,'$ORIGIN'
 */
#include <jni.h>
#include <stdint.h>
#include <stdlib.h>

#include "chaosultra.h"

#define NATIVE(ret, name) \
    JNIEXPORT ret JNICALL Java_cz_cuni_mff_cgg_teichmaa_chaosultra_hiprenderer_HipNative_##name

/* Status code -> exception class (the reference's classes where it has them). */
static const char *exception_class(int status) {
    switch (status) {
        case CU_ERR_ILLEGAL_ARGUMENT: return "java/lang/IllegalArgumentException";
        case CU_ERR_INITIALIZATION: return "cz/cuni/mff/cgg/teichmaa/chaosultra/cudarenderer/CudaInitializationException";
        case CU_ERR_RENDERER: return "cz/cuni/mff/cgg/teichmaa/chaosultra/rendering/FractalRendererException";
        case CU_ERR_ILLEGAL_STATE: return "java/lang/IllegalStateException";
        default: return "java/lang/RuntimeException";
    }
}

/* Returns 1 (and leaves a pending exception) when rc is not CU_OK. */
static int failed(JNIEnv *env, int rc) {
    if (rc == CU_OK) return 0;
    jclass cls = (*env)->FindClass(env, exception_class(rc));
    if (cls) (*env)->ThrowNew(env, cls, cu_last_error());
    return 1;
}

static void *model_ptr(JNIEnv *env, jobject buffer) {
    if ((*env)->GetDirectBufferCapacity(env, buffer) < (jlong)sizeof(cu_rendering_model)) {
        failed(env, CU_ERR_ILLEGAL_ARGUMENT);
        return NULL;
    }
    return (*env)->GetDirectBufferAddress(env, buffer);
}

NATIVE(jint, providerCreate)(JNIEnv *env, jclass cls, jstring dir) {
    (void)cls;
    const char *d = (*env)->GetStringUTFChars(env, dir, NULL);
    if (!d) return 0;
    cu_provider p = 0;
    int rc = cu_init(0);
    if (rc == CU_OK) rc = cu_provider_create(d, &p);
    (*env)->ReleaseStringUTFChars(env, dir, d);
    failed(env, rc);
    return p;
}

NATIVE(void, providerDestroy)(JNIEnv *env, jclass cls, jint p) {
    (void)cls;
    failed(env, cu_provider_destroy(p));
}

NATIVE(jstring, availableFractals)(JNIEnv *env, jclass cls, jint p) {
    (void)cls;
    char buf[1024];
    if (failed(env, cu_provider_available_fractals(p, buf, sizeof buf))) return NULL;
    return (*env)->NewStringUTF(env, buf);
}

NATIVE(jint, getRenderer)(JNIEnv *env, jclass cls, jint p, jstring name, jboolean force) {
    (void)cls;
    const char *n = (*env)->GetStringUTFChars(env, name, NULL);
    if (!n) return 0;
    cu_renderer r = 0;
    int rc = cu_provider_get_renderer(p, n, force ? 1 : 0, &r);
    (*env)->ReleaseStringUTFChars(env, name, n);
    failed(env, rc);
    return r;
}

NATIVE(void, initializeRendering)(JNIEnv *env, jclass cls, jint r, jint w, jint h, jintArray palette) {
    (void)cls;
    jsize len = (*env)->GetArrayLength(env, palette);
    jint *pal = (*env)->GetIntArrayElements(env, palette, NULL);
    if (!pal) return;
    int rc = cu_renderer_initialize_rendering(r, w, h, (const uint32_t *)pal, (int)len);
    (*env)->ReleaseIntArrayElements(env, palette, pal, JNI_ABORT);
    failed(env, rc);
}

NATIVE(void, freeRenderingResources)(JNIEnv *env, jclass cls, jint r) {
    (void)cls;
    failed(env, cu_renderer_free_rendering_resources(r));
}

NATIVE(jint, getState)(JNIEnv *env, jclass cls, jint r) {
    (void)cls;
    int s = 0;
    failed(env, cu_renderer_get_state(r, &s));
    return s;
}

NATIVE(jint, getWidth)(JNIEnv *env, jclass cls, jint r) {
    (void)cls;
    int w = 0;
    failed(env, cu_renderer_get_width(r, &w));
    return w;
}

NATIVE(jint, getHeight)(JNIEnv *env, jclass cls, jint r) {
    (void)cls;
    int h = 0;
    failed(env, cu_renderer_get_height(r, &h));
    return h;
}

NATIVE(void, launchDebugKernel)(JNIEnv *env, jclass cls, jint r) {
    (void)cls;
    failed(env, cu_renderer_launch_debug_kernel(r));
}

NATIVE(void, renderFast)(JNIEnv *env, jclass cls, jint r, jobject model) {
    (void)cls;
    cu_rendering_model *m = (cu_rendering_model *)model_ptr(env, model);
    if (m) failed(env, cu_renderer_render_fast(r, m));
}

NATIVE(void, renderQuality)(JNIEnv *env, jclass cls, jint r, jobject model) {
    (void)cls;
    cu_rendering_model *m = (cu_rendering_model *)model_ptr(env, model);
    if (m) failed(env, cu_renderer_render_quality(r, m));
}

NATIVE(void, close)(JNIEnv *env, jclass cls, jint r) {
    (void)cls;
    failed(env, cu_renderer_close(r));
}

NATIVE(void, setFractalCustomParams)(JNIEnv *env, jclass cls, jint r, jstring text) {
    (void)cls;
    const char *t = (*env)->GetStringUTFChars(env, text, NULL);
    if (!t) return;
    int rc = cu_renderer_set_fractal_custom_params(r, t);
    (*env)->ReleaseStringUTFChars(env, text, t);
    failed(env, rc);
}

NATIVE(jstring, getFractalName)(JNIEnv *env, jclass cls, jint r) {
    (void)cls;
    char buf[256];
    if (failed(env, cu_renderer_get_fractal_name(r, buf, sizeof buf))) return NULL;
    return (*env)->NewStringUTF(env, buf);
}

NATIVE(void, supplyDefaultValues)(JNIEnv *env, jclass cls, jint r, jobject model, jint w, jint h) {
    (void)cls;
    cu_rendering_model *m = (cu_rendering_model *)model_ptr(env, model);
    if (m) failed(env, cu_renderer_supply_default_values(r, m, w, h));
}

NATIVE(void, readOutput)(JNIEnv *env, jclass cls, jint r, jintArray rgba) {
    (void)cls;
    jsize len = (*env)->GetArrayLength(env, rgba);
    jint *px = (*env)->GetIntArrayElements(env, rgba, NULL);
    if (!px) return;
    int rc = cu_renderer_read_output(r, (uint32_t *)px, (size_t)len);
    (*env)->ReleaseIntArrayElements(env, rgba, px, rc == CU_OK ? 0 : JNI_ABORT);
    failed(env, rc);
}

NATIVE(jintArray, loadPaletteOrDefault)(JNIEnv *env, jclass cls, jstring path) {
    (void)cls;
    const char *p = path ? (*env)->GetStringUTFChars(env, path, NULL) : NULL;
    size_t cap = 1 << 16, count = 0;
    uint32_t *buf = (uint32_t *)malloc(cap * sizeof(uint32_t));
    int rc = buf ? cu_palette_load_or_default(p, buf, cap, &count, NULL) : CU_ERR_DEVICE;
    if (p) (*env)->ReleaseStringUTFChars(env, path, p);
    jintArray out = NULL;
    if (!failed(env, rc)) {
        out = (*env)->NewIntArray(env, (jsize)count);
        if (out) (*env)->SetIntArrayRegion(env, out, 0, (jsize)count, (const jint *)buf);
    }
    free(buf);
    return out;
}

NATIVE(void, saveImagePng)(JNIEnv *env, jclass cls, jstring path, jintArray rgba, jint w, jint h) {
    (void)cls;
    if ((*env)->GetArrayLength(env, rgba) < (jsize)w * h) {
        failed(env, CU_ERR_ILLEGAL_ARGUMENT);
        return;
    }
    const char *p = (*env)->GetStringUTFChars(env, path, NULL);
    if (!p) return;
    jint *px = (*env)->GetIntArrayElements(env, rgba, NULL);
    int rc = px ? cu_image_save_png(p, (const uint32_t *)px, w, h) : CU_ERR_DEVICE;
    if (px) (*env)->ReleaseIntArrayElements(env, rgba, px, JNI_ABORT);
    (*env)->ReleaseStringUTFChars(env, path, p);
    failed(env, rc);
}
