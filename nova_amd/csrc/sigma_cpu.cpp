// nova-mi355x: the sigma-approximation of the belief set on the host
// (src/pomdp/utilities/pomdp_sigma_cpu.cpp:50-118 in the reference).
//
// Each belief keeps its numDesiredNonZeroValues largest entries, renormalised by their sum
// sigma_b; sigma = min(1, min_b sigma_b). The entries are ordered by std::sort with a
// descending comparator on (value, state) pairs, as the reference orders them: that sort is not
// stable, and which of several equal values are kept is part of the result, so the same
// libstdc++ algorithm is used on the same input sequence. The reference reads past its arrays
// in two places; here the arrays hold r * numDesiredNonZeroValues entries, and entries past a
// row's rz count as value 0.
#include <algorithm>
#include <utility>
#include <vector>

#include <nova/pomdp/utilities/pomdp_sigma_cpu.h>

#include "nova_common.h"

namespace nova {

int pomdp_sigma_cpu(POMDP *pomdp, unsigned int numDesiredNonZeroValues, float &sigma)
{
    if (pomdp == nullptr || pomdp->z == 0 || pomdp->r == 0 || pomdp->rz == 0 ||
            pomdp->Z == nullptr || pomdp->B == nullptr || numDesiredNonZeroValues == 0) {
        NOVA_ERR("pomdp_sigma_cpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const unsigned r = pomdp->r, rz = pomdp->rz, k = numDesiredNonZeroValues;
    int *Znew = new int[(size_t)r * k];
    float *Bnew = new float[(size_t)r * k];
    sigma = 1.0f;
    std::vector<std::pair<float, int>> row;
    for (unsigned i = 0; i < r; i++) {
        row.clear();
        for (unsigned x = 0; x < rz; x++)
            row.emplace_back(pomdp->B[(size_t)i * rz + x], pomdp->Z[(size_t)i * rz + x]);
        std::sort(row.begin(), row.end(),
                  [](const std::pair<float, int> &l, const std::pair<float, int> &h) {
                      return l.first > h.first;
                  });
        float sigmab = 0.0f;
        for (unsigned x = 0; x < k && x < rz; x++)
            sigmab += row[x].first;
        if (sigmab < sigma)
            sigma = sigmab;
        for (unsigned x = 0; x < k; x++) {
            const bool keep = x < rz && row[x].first > 0.0f;
            Bnew[(size_t)i * k + x] = keep ? row[x].first / sigmab : 0.0f;
            Znew[(size_t)i * k + x] = keep ? row[x].second : -1;
        }
    }
    delete[] pomdp->Z;
    delete[] pomdp->B;
    pomdp->rz = k;
    pomdp->Z = Znew;
    pomdp->B = Bnew;
    return NOVA_SUCCESS;
}

} // namespace nova
