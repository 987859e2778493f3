package cz.cuni.mff.cgg.teichmaa.chaosultra.hiprenderer;

import java.nio.ByteBuffer;

/**
 * JNI entry points onto libchaosultra_hip (include/chaosultra.h): one native
 * method per C-ABI call of the renderer layer.  A failing call throws the
 * exception class its status code stands for (IllegalArgumentException,
 * CudaInitializationException, FractalRendererException,
 * IllegalStateException, or RuntimeException for any other device error),
 * with the library's message (chaosultra_jni.c).
 */
final class HipNative {
    static {
        System.loadLibrary("chaosultra_jni");  // links libchaosultra_hip.so
    }

    private HipNative() {
    }

    /** sizeof(cu_rendering_model): the layout ModelMarshal writes. */
    static final int MODEL_BYTES = 600;

    static native int providerCreate(String kernelsDir);
    static native void providerDestroy(int provider);
    static native String availableFractals(int provider);
    static native int getRenderer(int provider, String fractalName, boolean forceReload);

    static native void initializeRendering(int renderer, int width, int height, int[] paletteRgba);
    static native void freeRenderingResources(int renderer);
    static native int getState(int renderer);
    static native int getWidth(int renderer);
    static native int getHeight(int renderer);
    static native void launchDebugKernel(int renderer);
    static native void renderFast(int renderer, ByteBuffer model);
    static native void renderQuality(int renderer, ByteBuffer model);
    static native void close(int renderer);
    static native void setFractalCustomParams(int renderer, String text);
    static native String getFractalName(int renderer);
    static native void supplyDefaultValues(int renderer, ByteBuffer model, int canvasWidth, int canvasHeight);
    /** The composed RGBA8 image, width*height words, R in the least significant byte. */
    static native void readOutput(int renderer, int[] rgba);

    /** ImageHelpers.loadColorPaletteOrDefault on the native side. */
    static native int[] loadPaletteOrDefault(String path);
    /** ImageHelpers.saveImageToFile(rgba, w, h, path, "png") on the native side. */
    static native void saveImagePng(String path, int[] rgba, int width, int height);
}
