// plife_jni.c -- JNI glue of com.particle_life.backend.GpuPhysics onto the plife C ABI (include/plife.h).
// Build (needs a JDK; not built in this repository): make -C particle-life-app_amd/java JAVA_HOME=...
#include <stddef.h>
#include <stdint.h>
#include <jni.h>
#include "plife.h"

static void throw_status(JNIEnv *env, int rc) {
    if (rc >= 0) return;
    const char *cls = rc == PLIFE_ERR_ARGUMENT ? "java/lang/IllegalArgumentException"
                                               : "java/lang/RuntimeException";
    (*env)->ThrowNew(env, (*env)->FindClass(env, cls), plife_last_error());
}

JNIEXPORT jlong JNICALL Java_com_particle_1life_backend_GpuPhysics_nCreate(JNIEnv *env, jclass c, jint dev) {
    plife_physics *h = 0;
    throw_status(env, plife_create(dev, &h));
    return (jlong)(intptr_t)h;
}
JNIEXPORT void JNICALL Java_com_particle_1life_backend_GpuPhysics_nDestroy(JNIEnv *env, jclass c, jlong h) {
    plife_destroy((plife_physics *)(intptr_t)h);
}
JNIEXPORT void JNICALL Java_com_particle_1life_backend_GpuPhysics_nSetSettings(JNIEnv *env, jclass c, jlong h,
        jint dim, jint wrap, jdouble rmax, jdouble friction, jdouble force, jdouble dt, jdouble beta,
        jint ntypes, jdoubleArray m) {
    plife_settings s = {dim, wrap, rmax, friction, force, dt, beta};
    jdouble *mm = (*env)->GetDoubleArrayElements(env, m, 0);
    int rc = plife_set_settings((plife_physics *)(intptr_t)h, &s, ntypes, mm);
    (*env)->ReleaseDoubleArrayElements(env, m, mm, JNI_ABORT);
    throw_status(env, rc);
}
JNIEXPORT void JNICALL Java_com_particle_1life_backend_GpuPhysics_nSetParticles(JNIEnv *env, jclass c, jlong h,
        jint n, jdoubleArray pos, jdoubleArray vel, jintArray types) {
    jdouble *p = (*env)->GetDoubleArrayElements(env, pos, 0), *v = (*env)->GetDoubleArrayElements(env, vel, 0);
    jint *t = (*env)->GetIntArrayElements(env, types, 0);
    int rc = plife_set_particles((plife_physics *)(intptr_t)h, n, p, v, t);
    (*env)->ReleaseDoubleArrayElements(env, pos, p, JNI_ABORT);
    (*env)->ReleaseDoubleArrayElements(env, vel, v, JNI_ABORT);
    (*env)->ReleaseIntArrayElements(env, types, t, JNI_ABORT);
    throw_status(env, rc);
}
JNIEXPORT void JNICALL Java_com_particle_1life_backend_GpuPhysics_nGetParticles(JNIEnv *env, jclass c, jlong h,
        jint n, jdoubleArray pos, jdoubleArray vel, jintArray types, jintArray ids) {
    /* any array may be null: that part is not downloaded */
    jdouble *p = pos ? (*env)->GetDoubleArrayElements(env, pos, 0) : NULL;
    jdouble *v = vel ? (*env)->GetDoubleArrayElements(env, vel, 0) : NULL;
    jint *t = types ? (*env)->GetIntArrayElements(env, types, 0) : NULL;
    jint *i = ids ? (*env)->GetIntArrayElements(env, ids, 0) : NULL;
    int rc = plife_get_particles((plife_physics *)(intptr_t)h, n, p, v, t, i);
    if (p) (*env)->ReleaseDoubleArrayElements(env, pos, p, 0);
    if (v) (*env)->ReleaseDoubleArrayElements(env, vel, v, 0);
    if (t) (*env)->ReleaseIntArrayElements(env, types, t, 0);
    if (i) (*env)->ReleaseIntArrayElements(env, ids, i, 0);
    throw_status(env, rc);
}
JNIEXPORT void JNICALL Java_com_particle_1life_backend_GpuPhysics_nSetParticleSlots(JNIEnv *env, jclass c, jlong h,
        jint k, jintArray slots, jdoubleArray pos, jdoubleArray vel, jintArray types) {
    jint *s = (*env)->GetIntArrayElements(env, slots, 0), *t = (*env)->GetIntArrayElements(env, types, 0);
    jdouble *p = (*env)->GetDoubleArrayElements(env, pos, 0), *v = (*env)->GetDoubleArrayElements(env, vel, 0);
    int rc = plife_set_particle_slots((plife_physics *)(intptr_t)h, k, s, p, v, t);
    (*env)->ReleaseIntArrayElements(env, slots, s, JNI_ABORT);
    (*env)->ReleaseIntArrayElements(env, types, t, JNI_ABORT);
    (*env)->ReleaseDoubleArrayElements(env, pos, p, JNI_ABORT);
    (*env)->ReleaseDoubleArrayElements(env, vel, v, JNI_ABORT);
    throw_status(env, rc);
}
JNIEXPORT void JNICALL Java_com_particle_1life_backend_GpuPhysics_nUpdate(JNIEnv *env, jclass c, jlong h, jint k) {
    throw_status(env, plife_update((plife_physics *)(intptr_t)h, k));
}
JNIEXPORT void JNICALL Java_com_particle_1life_backend_GpuPhysics_nForceUpdateStop(JNIEnv *env, jclass c, jlong h) {
    plife_force_update_stop((plife_physics *)(intptr_t)h);
}
JNIEXPORT void JNICALL Java_com_particle_1life_backend_GpuPhysics_nSetPositions(JNIEnv *env, jclass c, jlong h,
        jint preset, jlong seed) {
    throw_status(env, plife_set_positions((plife_physics *)(intptr_t)h, preset, (uint64_t)seed));
}
JNIEXPORT void JNICALL Java_com_particle_1life_backend_GpuPhysics_nSetTypes(JNIEnv *env, jclass c, jlong h,
        jint preset, jlong seed) {
    throw_status(env, plife_set_types((plife_physics *)(intptr_t)h, preset, (uint64_t)seed));
}
