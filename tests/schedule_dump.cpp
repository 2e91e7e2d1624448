// Prints the banded launch schedules of a grid, for tests/test_schedule.py:
//   schedule_dump NTILES IZ CUS FILL
// prints, per candidate, "cand <bands> <units> <span> <deal>", its band
// thicknesses ("th t0 t1 ..."), then one "u <tile> <zb> <ze>" line per unit in
// launch order.
#include <cstdio>
#include <cstdlib>

#include "../multicomponent_color_gradient_amd/csrc/schedule.hpp"

int main(int argc, char **argv) {
    if (argc != 5) return 2;
    const int ntiles = std::atoi(argv[1]), IZ = std::atoi(argv[2]), cus = std::atoi(argv[3]);
    const double fill = std::atof(argv[4]);
    for (const mcg::BandedSchedule &b : mcg::banded_rules(ntiles, IZ, cus, fill)) {
        std::printf("cand %d %zu %.3f %d\nth", b.bands, b.units.size(), mcg::schedule_span(b.units, cus, fill), int(b.deal));
        for (int t : b.thickness) std::printf(" %d", t);
        std::printf("\n");
        for (const mcg::FusedUnit &u : b.units) std::printf("u %u %u %u\n", u.tile, u.zb, u.ze);
    }
    return 0;
}
