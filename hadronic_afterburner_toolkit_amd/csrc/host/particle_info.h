// particle_info -- the particle record of the reference (src/particle_info.h:5-12).
#ifndef HBT_AMD_PARTICLE_INFO_H
#define HBT_AMD_PARTICLE_INFO_H

typedef struct {
    int monval;
    double mass;
    int baryon;
    double px, py, pz, E;
    double x, y, z, t;
    double rap_y, rap_eta, pT, phi_p;
} particle_info;

#endif
