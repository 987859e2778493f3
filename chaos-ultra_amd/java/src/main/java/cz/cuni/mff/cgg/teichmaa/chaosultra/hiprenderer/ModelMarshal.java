package cz.cuni.mff.cgg.teichmaa.chaosultra.hiprenderer;

import cz.cuni.mff.cgg.teichmaa.chaosultra.rendering.model.DefaultFractalModel;
import cz.cuni.mff.cgg.teichmaa.chaosultra.rendering.model.PlaneSegment;
import cz.cuni.mff.cgg.teichmaa.chaosultra.rendering.model.RenderingModel;
import cz.cuni.mff.cgg.teichmaa.chaosultra.util.FloatPrecision;

import java.nio.ByteBuffer;
import java.nio.ByteOrder;
import java.nio.charset.StandardCharsets;

/**
 * Copies a RenderingModel into a direct buffer laid out as cu_rendering_model
 * (include/chaosultra.h) and writes the renderer's outputs back into the model,
 * as the Java setters of the reference's renderer would have.
 */
final class ModelMarshal {
    private static final int LBX = 0, LBY = 8, RTX = 16, RTY = 24;
    private static final int MAX_ITER = 32, MAX_SS = 36, ADAPTIVE = 40, VISUALISE = 44;
    private static final int FOVEATION = 48, FOCUS_X = 52, FOCUS_Y = 56, ZOOMING = 60, ZOOMING_IN = 64;
    private static final int REUSE = 68, REUSE_DIRTY = 72, PRECISION = 76, ERROR_COUNT = 80;
    private static final int LAST_ERROR = 84, CUSTOM_PARAMS = 340, TEXT_BYTES = 256;

    private ModelMarshal() {
    }

    static ByteBuffer allocate() {
        return ByteBuffer.allocateDirect(HipNative.MODEL_BYTES).order(ByteOrder.nativeOrder());
    }

    static void write(RenderingModel m, ByteBuffer b) {
        PlaneSegment s = m.getPlaneSegment();
        b.putDouble(LBX, s.getLeftBottom().getX());
        b.putDouble(LBY, s.getLeftBottom().getY());
        b.putDouble(RTX, s.getRightTop().getX());
        b.putDouble(RTY, s.getRightTop().getY());
        b.putInt(MAX_ITER, m.getMaxIterations());
        b.putFloat(MAX_SS, m.getMaxSuperSampling());
        b.putInt(ADAPTIVE, m.isUseAdaptiveSuperSampling() ? 1 : 0);
        b.putInt(VISUALISE, m.isVisualiseSampleCount() ? 1 : 0);
        b.putInt(FOVEATION, m.isUseFoveatedRendering() ? 1 : 0);
        b.putInt(FOCUS_X, m.getMouseFocus().getX());
        b.putInt(FOCUS_Y, m.getMouseFocus().getY());
        b.putInt(ZOOMING, m.isZooming() ? 1 : 0);
        b.putInt(ZOOMING_IN, m.isZoomingIn() ? 1 : 0);
        b.putInt(REUSE, m.isUseSampleReuse() ? 1 : 0);
        b.putInt(REUSE_DIRTY, m.isSampleReuseCacheDirty() ? 1 : 0);
        b.putInt(PRECISION, m.getFloatingPointPrecision().ordinal());
        b.putInt(ERROR_COUNT, 0);
        b.put(LAST_ERROR, (byte) 0);
        putText(b, CUSTOM_PARAMS, m instanceof DefaultFractalModel
                ? ((DefaultFractalModel) m).getFractalCustomParams() : "");
    }

    /** After renderFast / renderQuality: reuse-cache flag, precision, logged errors. */
    static void readBack(ByteBuffer b, RenderingModel m) {
        m.setSampleReuseCacheDirty(b.getInt(REUSE_DIRTY) != 0);
        m.setFloatingPointPrecision(FloatPrecision.values()[b.getInt(PRECISION)]);
        if (b.getInt(ERROR_COUNT) > 0)
            m.logError(getText(b, LAST_ERROR));  // the last of the render's logged errors
    }

    /**
     * After supplyDefaultValues: the fields a module supplies.  The module gives a
     * centre and a zoom (the segment height); the model maps them onto its own
     * canvas with setPlaneSegmentFromCenter, as the reference's modules do.
     */
    static void readDefaults(ByteBuffer b, DefaultFractalModel m) {
        double lbx = b.getDouble(LBX), lby = b.getDouble(LBY), rtx = b.getDouble(RTX), rty = b.getDouble(RTY);
        m.setPlaneSegmentFromCenter((lbx + rtx) / 2, (lby + rty) / 2, rty - lby);
        m.setMaxIterations(b.getInt(MAX_ITER));
        m.setMaxSuperSampling(b.getFloat(MAX_SS));
        m.setFractalCustomParams(getText(b, CUSTOM_PARAMS));
    }

    private static void putText(ByteBuffer b, int offset, String s) {
        byte[] bytes = (s == null ? "" : s).getBytes(StandardCharsets.UTF_8);
        int n = Math.min(bytes.length, TEXT_BYTES - 1);
        for (int i = 0; i < n; i++) b.put(offset + i, bytes[i]);
        b.put(offset + n, (byte) 0);
    }

    private static String getText(ByteBuffer b, int offset) {
        int n = 0;
        while (n < TEXT_BYTES && b.get(offset + n) != 0) n++;
        byte[] bytes = new byte[n];
        for (int i = 0; i < n; i++) bytes[i] = b.get(offset + i);
        return new String(bytes, StandardCharsets.UTF_8);
    }
}
