// ref_ladder.cpp -- TEST INFRASTRUCTURE: exposes the reference's own
// temperature ladder, Loader::set_temperature (src/frontend/frontend.cpp:222-260),
// compiled unmodified from its source by oracle/Makefile into oracle/_ref/.
// The ladder is a protected member of the reference's Loader, so a derived
// class sets the ladder parameters the CSV line would (ntemp, temp_high,
// temp_low; frontend.cpp:104-106) and calls it.  Only the Loader's
// constructor, destructor and set_temperature are linked; the link drops
// the rest of frontend.cpp (its loaders need the Fortran u3b routine).
#include <cstring>

#include <geauxdock.h>
#include <size.h>
#include <frontend.h>

namespace {
struct LadderProbe : Loader {
    int run(int n, float t_low, float t_high, float *minus_beta, int *order)
    {
        if (n < 1 || n > MAX_TMP) return -1;
        exchgpara.num_temp = n;
        exchgpara.floor_temp = t_low;
        exchgpara.ceiling_temp = t_high;
        set_temperature();
        for (int i = 0; i < n; ++i) {
            minus_beta[i] = temp[i].minus_beta;
            order[i] = temp[i].order;
        }
        return 0;
    }
};
} // namespace

extern "C" __attribute__((visibility("default"))) int gdref_set_temperature(int n, float t_low, float t_high,
                                                                              float *minus_beta, int *order)
{
    LadderProbe p;
    return p.run(n, t_low, t_high, minus_beta, order);
}
