package cz.cuni.mff.cgg.teichmaa.chaosultra.hiprenderer;

import cz.cuni.mff.cgg.teichmaa.chaosultra.rendering.FractalRenderer;
import cz.cuni.mff.cgg.teichmaa.chaosultra.rendering.FractalRendererProvider;

import java.util.Arrays;
import java.util.Set;
import java.util.TreeSet;

/**
 * FractalRendererProvider on MI355X: the native provider (cu_provider_*) loads
 * one gfx950 code object per fractal from the kernels directory, as
 * CudaFractalRendererProvider loads one PTX module per fractal from
 * {@code cudaKernelsDir}.  Drop-in for
 * cudarenderer/CudaFractalRendererProvider.java: same methods, same renderer
 * lifecycle (getRenderer closes the active renderer when it switches fractal
 * or when forceReload is set).
 */
public class HipFractalRendererProvider implements FractalRendererProvider {
    /** Same system property as the reference; its value names the .hsaco directory. */
    public static final String KERNELS_DIR_PROPERTY = "cudaKernelsDir";

    private final int handle;

    public HipFractalRendererProvider() {
        this(System.getProperty(KERNELS_DIR_PROPERTY, "chaos-ultra_amd/kernels"));
    }

    public HipFractalRendererProvider(String kernelsDir) {
        handle = HipNative.providerCreate(kernelsDir);
    }

    @Override
    public FractalRenderer getDefaultRenderer() {
        return getRenderer("mandelbrot", false);
    }

    @Override
    public FractalRenderer getRenderer(String fractalName, boolean forceReload) {
        return new HipFractalRenderer(HipNative.getRenderer(handle, fractalName, forceReload));
    }

    @Override
    public Set<String> getAvailableFractals() {
        return new TreeSet<>(Arrays.asList(HipNative.availableFractals(handle).split(",")));
    }
}
