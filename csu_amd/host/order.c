/*
 * order.c -- cell renumbering support for csu_amd/renumber.py: the greedy
 * colouring pass, the one sequential step of the colour/RCM order.
 *
 * Cells are visited in the given order (reverse Cuthill-McKee); each takes
 * the smallest colour none of its already coloured neighbours has.  The
 * LU-SGS sweep of a mesh renumbered colour-major then has as many levels as
 * colours (implicit.c:274-482 sweep cells in index order).
 */
#include <stdlib.h>
#include <string.h>
#include "csu_host.h"

int csu_greedy_colour(int32_t n, const int64_t* indptr, const int32_t* indices, const int64_t* order,
                      int32_t* colour)
{
    if (n < 0 || (n > 0 && (!indptr || !indices || !order || !colour))) return CSU_ERR_ARG;
    for (int32_t i = 0; i < n; i++) colour[i] = -1;
    int32_t cap = 64;
    unsigned char* used = calloc((size_t)cap, 1);
    if (!used) return CSU_ERR_NOMEM;
    for (int32_t t = 0; t < n; t++) {
        const int64_t i = order[t];
        if (i < 0 || i >= n) { free(used); return CSU_ERR_ARG; }
        int32_t deg = (int32_t)(indptr[i + 1] - indptr[i]);
        if (deg + 1 > cap) {
            cap = 2 * (deg + 1);
            free(used);
            used = calloc((size_t)cap, 1);
            if (!used) return CSU_ERR_NOMEM;
        }
        /* colours above deg cannot block the smallest free one */
        for (int64_t q = indptr[i]; q < indptr[i + 1]; q++) {
            int32_t c = colour[indices[q]];
            if (c >= 0 && c <= deg) used[c] = 1;
        }
        int32_t k = 0;
        while (used[k]) k++;
        colour[i] = k;
        for (int64_t q = indptr[i]; q < indptr[i + 1]; q++) {
            int32_t c = colour[indices[q]];
            if (c >= 0 && c <= deg) used[c] = 0;
        }
    }
    free(used);
    return CSU_OK;
}
