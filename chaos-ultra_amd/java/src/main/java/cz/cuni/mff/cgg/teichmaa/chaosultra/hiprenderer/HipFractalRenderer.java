package cz.cuni.mff.cgg.teichmaa.chaosultra.hiprenderer;

import com.jogamp.opengl.GL;
import com.jogamp.opengl.GL2;
import com.jogamp.opengl.GLContext;
import cz.cuni.mff.cgg.teichmaa.chaosultra.rendering.FractalRenderer;
import cz.cuni.mff.cgg.teichmaa.chaosultra.rendering.FractalRendererState;
import cz.cuni.mff.cgg.teichmaa.chaosultra.rendering.model.DefaultFractalModel;
import cz.cuni.mff.cgg.teichmaa.chaosultra.rendering.model.GLParams;
import cz.cuni.mff.cgg.teichmaa.chaosultra.rendering.model.GLTexture;
import cz.cuni.mff.cgg.teichmaa.chaosultra.rendering.model.RenderingModel;

import java.nio.ByteBuffer;
import java.nio.IntBuffer;

/**
 * FractalRenderer on MI355X.  Every method forwards to one native call of the
 * renderer layer (cu_renderer_*), which runs the reference's kernels rebuilt
 * for gfx950 on HIP.  Drop-in for cudarenderer/CudaFractalRenderer.java.
 * <p>
 * There is no CUDA-GL interop on the HIP side of a compute GPU: the palette is
 * read from its GL texture once in initializeRendering, and after each render
 * the composed RGBA8 image is read back and uploaded to the output texture
 * (glTexSubImage2D) — one PCIe round trip per frame, where the reference wrote
 * the texture from the kernel through a mapped surface.  Like the reference,
 * these calls need the GL context current (the AWT event thread).
 */
public class HipFractalRenderer implements FractalRenderer {
    private final int handle;
    private final ByteBuffer model = ModelMarshal.allocate();
    private GLTexture output;
    private int[] rgba = new int[0];

    HipFractalRenderer(int handle) {
        this.handle = handle;
    }

    @Override
    public void initializeRendering(GLParams glParams) {
        output = glParams.getOutput();
        GLTexture palette = glParams.getPalette();
        int[] paletteRgba = new int[glParams.getPaletteLength()];
        GL2 gl = GLContext.getCurrentGL().getGL2();
        gl.glBindTexture(palette.getTarget(), palette.getHandle().getValue());
        gl.glGetTexImage(palette.getTarget(), 0, GL.GL_RGBA, GL.GL_UNSIGNED_BYTE, IntBuffer.wrap(paletteRgba));
        HipNative.initializeRendering(handle, output.getWidth(), output.getHeight(), paletteRgba);
        rgba = new int[output.getWidth() * output.getHeight()];
    }

    @Override
    public void freeRenderingResources() {
        HipNative.freeRenderingResources(handle);
    }

    @Override
    public FractalRendererState getState() {
        return FractalRendererState.values()[HipNative.getState(handle)];
    }

    @Override
    public int getWidth() {
        return HipNative.getWidth(handle);
    }

    @Override
    public int getHeight() {
        return HipNative.getHeight(handle);
    }

    @Override
    public void launchDebugKernel() {
        HipNative.launchDebugKernel(handle);
    }

    @Override
    public void renderFast(RenderingModel m) {
        ModelMarshal.write(m, model);
        HipNative.renderFast(handle, model);
        ModelMarshal.readBack(model, m);
        upload();
    }

    @Override
    public void renderQuality(RenderingModel m) {
        ModelMarshal.write(m, model);
        HipNative.renderQuality(handle, model);
        ModelMarshal.readBack(model, m);
        upload();
    }

    @Override
    public void close() {
        HipNative.close(handle);
    }

    @Override
    public void setFractalCustomParams(String text) {
        HipNative.setFractalCustomParams(handle, text);
    }

    @Override
    public String getFractalName() {
        return HipNative.getFractalName(handle);
    }

    @Override
    public void supplyDefaultValues(DefaultFractalModel m) {
        ModelMarshal.write(m, model);
        HipNative.supplyDefaultValues(handle, model, 1, 1);  // centre and zoom do not depend on the canvas
        ModelMarshal.readDefaults(model, m);
    }

    /** The rendered image into the output texture. */
    private void upload() {
        HipNative.readOutput(handle, rgba);
        GL2 gl = GLContext.getCurrentGL().getGL2();
        gl.glBindTexture(output.getTarget(), output.getHandle().getValue());
        gl.glTexSubImage2D(output.getTarget(), 0, 0, 0, output.getWidth(), output.getHeight(),
                GL.GL_RGBA, GL.GL_UNSIGNED_BYTE, IntBuffer.wrap(rgba));
    }
}
