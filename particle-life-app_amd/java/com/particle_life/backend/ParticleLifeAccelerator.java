// ParticleLifeAccelerator -- the app's accelerator lambda (app/Main.java:275-280) as a named type, so
// that GpuPhysics can read beta and run the same law on the device.
// Not compiled in this repository (no JDK in its build image); see INTEGRATION.md.
package com.particle_life.backend;
import org.joml.Vector3d;

/** The app's lambda (app/Main.java:275-280) as a named type, so the GPU can read beta. */
public final class ParticleLifeAccelerator implements Accelerator {
    public final double beta;
    public ParticleLifeAccelerator(double beta) { this.beta = beta; }
    @Override public Vector3d accelerate(double a, Vector3d pos) {
        double dist = pos.length();
        double force = dist < beta ? (dist / beta - 1) : a * (1 - Math.abs(1 + beta - 2 * dist) / (1 - beta));
        return pos.mul(force / dist);
    }
}
