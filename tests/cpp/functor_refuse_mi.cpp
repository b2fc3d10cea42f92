// A functor without a device implementation is refused as the reference refuses what it
// cannot do: a message and an abnormal exit (particles/LATfield2_Particles.hpp:217-244
// takes any host function; the device path runs the ones it implements).  The test expects
// exit status 255 and the message on stderr.
#include "LATfield2.hpp"

using namespace LATfield2;

// a user's own velocity update with the reference's functor signature
Real my_update(double dtau, double, part_simple *p, double *, part_simple_info, Field<Real> **, Site *, int, double *,
               double *, int) {
    for (int i = 0; i < 3; ++i) p->vel[i] *= (Real)(1.0 - dtau);
    return 0;
}

int main() {
    parallel.initialize(1, 1);
    const int N = 8;
    Lattice lat_part(3, N, 0);
    part_simple_info info;
    info.mass = 1.0;
    set_parts_typename(&info, "part_simple");
    part_simple_dataType dtype;
    Real box[3] = {1, 1, 1};
    Particles<part_simple, part_simple_info, part_simple_dataType> parts;
    parts.initialize(info, dtype, &lat_part, box);
    part_simple p;
    p.ID = 0;
    for (int d = 0; d < 3; ++d) {
        p.pos[d] = 0.3;
        p.vel[d] = 1.0;
    }
    parts.addParticle_global(p);
    parts.updateVel(my_update, 0.1);
    COUT << "not reached" << endl;
    return 0;
}
