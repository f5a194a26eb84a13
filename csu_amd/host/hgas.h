/*
 * hgas.h -- calorically / thermally perfect gas on the host (gasprop.c:14-187):
 * the free-stream state of case.c and the derived columns of solution2.csv.
 */
#ifndef CSU_HGAS_H
#define CSU_HGAS_H
#include <math.h>

typedef struct { double gamma, R, Cp, Cv, cc[6]; int tp; } hgas;

static inline void hgas_init(hgas* g, double gamma, double R, int tp)
{
    g->gamma = gamma; g->R = R;
    g->Cp = g->gamma * g->R / (g->gamma - 1);
    g->Cv = g->Cp - g->R;
    g->tp = tp;
    g->cc[0] = 1.81716e3;
    g->cc[1] = 9.890504e2 - 287.0530;
    g->cc[2] = -9.595592e-3;
    g->cc[3] = 1.041469e-4;
    g->cc[4] = -4.433065e-8;
    g->cc[5] = 5.879263e-12;
}

static inline double hgas_T2e(const hgas* g, double T)
{
    if (!g->tp) return g->Cv * T;
    double P = 1, aux = g->cc[0];
    for (int i = 1; i < 6; i++) { P *= T; aux += g->cc[i] * P; }
    return aux;
}

static inline double hgas_T2Cv(const hgas* g, double T)
{
    if (!g->tp) return g->Cv;
    double P = 1, aux = g->cc[1];
    for (int i = 2; i < 6; i++) { P *= T; aux += i * g->cc[i] * P; }
    return aux;
}

static inline double hgas_T2c(const hgas* g, double T)
{
    double Cv = hgas_T2Cv(g, T);
    double gam = (Cv + g->R) / Cv;
    return sqrt(gam * g->R * T);
}

/* gasprop_T2fentro / gasprop_Tp2entropy (gasprop.c:161-187) */
static inline double hgas_Tp2entropy(const hgas* g, double T, double p)
{
    double aux;
    if (g->tp) {
        double P = 1;
        aux = (g->cc[1] + g->R) * log(T);
        for (int i = 2; i < 6; i++) {
            P *= T;
            aux += i * g->cc[i] * P / (i - 1);
        }
    } else {
        aux = g->Cp * log(T);
    }
    return aux - g->R * log(p);
}

#endif
