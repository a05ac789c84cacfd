// GpuPhysics -- Physics.update() on an MI355X through the plife C ABI (include/plife.h), by JNI.
// Drop-in for backend/Physics.java: re-parent app/ExtendedPhysics.java (`extends Physics`, line 13)
// onto GpuPhysics and construct it with a ParticleLifeAccelerator (app/Main.java:275-285).
// Not compiled in this repository (no JDK in its build image); see INTEGRATION.md.
package com.particle_life.backend;

public class GpuPhysics extends Physics {
    static { System.loadLibrary("plife_jni"); }           // libplife_jni.so links libplife.so

    private long handle;                                   // plife_physics*
    private boolean hostDirty = true;                      // host particles may differ from device
    private Particle[] byId;                               // `particles` at upload: device ids index it

    public GpuPhysics(Accelerator accelerator, PositionSetter ps, MatrixGenerator mg, TypeSetter ts) {
        super(accelerator, ps, mg, ts);
        if (!(accelerator instanceof ParticleLifeAccelerator))
            throw new UnsupportedOperationException("GpuPhysics needs a ParticleLifeAccelerator");
        handle = nCreate(0);
    }

    /** Physics.update() (Physics.java:112-134) on the device. */
    @Override public void update() {
        pushSettings();
        if (hostDirty) { upload(); hostDirty = false; }
        nUpdate(handle, 1);
        download();          // keeps `particles` the API-visible truth, in cell order like the reference
    }

    @Override public void forceUpdateStop() { nForceUpdateStop(handle); }   // any thread
    @Override public void kill() { if (handle != 0) { nDestroy(handle); handle = 0; } }

    /** Call after editing `particles`, the matrix size or types between steps (cursors, loadState). */
    public void markParticlesDirty() { hostDirty = true; }

    /** A cursor tool that changed a few particles (indices into `particles`) sends only those. */
    public void editParticles(int[] indices) {
        if (hostDirty) return;                               // everything goes up at the next update
        int k = indices.length;
        double[] pos = new double[3 * k], vel = new double[3 * k];
        int[] types = new int[k];
        for (int i = 0; i < k; i++) {
            Particle p = particles[indices[i]];
            pos[3 * i] = p.position.x; pos[3 * i + 1] = p.position.y; pos[3 * i + 2] = p.position.z;
            vel[3 * i] = p.velocity.x; vel[3 * i + 1] = p.velocity.y; vel[3 * i + 2] = p.velocity.z;
            types[i] = p.type;
        }
        nSetParticleSlots(handle, k, indices, pos, vel, types);   // `particles` is in device order
    }

    private void pushSettings() {
        int n = settings.matrix.size();
        double[] m = new double[n * n];
        for (int i = 0; i < n; i++) for (int j = 0; j < n; j++) m[i * n + j] = settings.matrix.get(i, j);
        nSetSettings(handle, 2, settings.wrap ? 1 : 0, settings.rmax, settings.friction,
                     settings.force, settings.dt, ((ParticleLifeAccelerator) accelerator).beta, n, m);
    }

    private void upload() {
        int n = particles.length;
        double[] pos = new double[3 * n], vel = new double[3 * n];
        int[] types = new int[n];
        for (int i = 0; i < n; i++) {
            Particle p = particles[i];
            pos[3 * i] = p.position.x; pos[3 * i + 1] = p.position.y; pos[3 * i + 2] = p.position.z;
            vel[3 * i] = p.velocity.x; vel[3 * i + 1] = p.velocity.y; vel[3 * i + 2] = p.velocity.z;
            types[i] = p.type;
        }
        nSetParticles(handle, n, pos, vel, types);
        byId = particles.clone();
    }

    private void download() {
        int n = particles.length;
        double[] pos = new double[3 * n], vel = new double[3 * n];
        int[] types = new int[n], ids = new int[n];
        nGetParticles(handle, n, pos, vel, types, ids);
        Particle[] sorted = new Particle[n];
        for (int i = 0; i < n; i++) {
            Particle p = byId[ids[i]];                       // ids = indices at upload time
            p.position.set(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
            p.velocity.set(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]);
            p.type = types[i];
            sorted[i] = p;
        }
        particles = sorted;  // device order, as the reference's array after its sort
    }

    /** PositionSetterProvider's preset `preset` (its list order) run on the device over all particles. */
    public void setPositionsPreset(int preset, long seed) {
        pushSettings();
        if (hostDirty) { upload(); hostDirty = false; }
        nSetPositions(handle, preset, seed);
        download();
    }

    /** TypeSetterProvider's preset `preset` (its list order) run on the device over all particles. */
    public void setTypesPreset(int preset, long seed) {
        pushSettings();
        if (hostDirty) { upload(); hostDirty = false; }
        nSetTypes(handle, preset, seed);
        download();
    }

    private static native long nCreate(int device);
    private static native void nDestroy(long h);
    private static native void nSetSettings(long h, int dim, int wrap, double rmax, double friction,
                                            double force, double dt, double beta, int ntypes, double[] m);
    private static native void nSetParticles(long h, int n, double[] pos, double[] vel, int[] types);
    private static native void nGetParticles(long h, int n, double[] pos, double[] vel, int[] types, int[] ids);
    private static native void nSetParticleSlots(long h, int k, int[] slots, double[] pos, double[] vel, int[] types);
    private static native void nUpdate(long h, int nsteps);
    private static native void nForceUpdateStop(long h);
    private static native void nSetPositions(long h, int preset, long seed);
    private static native void nSetTypes(long h, int preset, long seed);
}
