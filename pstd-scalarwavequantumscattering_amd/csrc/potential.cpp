// Default potential hook linked into the qscat_amd driver, the user plugin
// point of the reference (potential.cpp:31-42, run_environment.h:82):
// scaled, centre-relative coordinates in; energy in E0 units out.  This
// example is a spherical barrier of height 5% of the example energy
// (0.327216805) and radius 2 pi (one incident wavelength).  Replace this file
// (or pass another hook to pstd_create) to model another potential; the hook
// must be pure and reentrant.
#include <cmath>

extern "C" double pstd_default_potential(double x, double y, double z) {
    const double V0 = 0.327216805 * 0.05;
    const double R = 6.283185307179586476925286766559005768394;
    return (x * x + y * y + z * z <= R * R) ? V0 : 0.0;
}
