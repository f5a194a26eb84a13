/*
 * mesh.c -- SU2 reader and connectivity in the reference's order.
 *
 * The device kernels sum each cell's face fluxes and gradient terms in the
 * order the reference does, so the connectivity must come out in that order:
 *
 *   - tokens split on ' ', '\r', '\n' (reference mesh.c:10-42), read in the
 *     fixed sequence of mesh.c:115-220 / 79-113;
 *   - cells with negative (axisymmetric-weighted) area get their vertex list
 *     reversed (mesh.c:222-242);
 *   - interior faces are enumerated as the reference's hash table lists them
 *     (mesh.c:1083-1090, 1172-1218): bucket floor(S*frac((p0+p1)*c)) with
 *     S = Nelem+Npoin and c = (sqrt(5)-1)/2, then order of first insertion
 *     inside a bucket.  Here that order is produced by a sort instead of a
 *     chained hash table: faces are paired by sorting the directed edges on
 *     their undirected key, then sorted on (bucket, first insertion);
 *   - neighbour lists: interior neighbours in face order, then boundary faces
 *     in marker/file order (mesh.c:792-830); a boundary face attaches to the
 *     lowest-index cell holding both its points (mesh.c:812-830);
 *   - boundary faces are turned CCW with respect to their cell
 *     (mesh.c:948-1009); per-cell signed face ids follow face order
 *     (mesh.c:1011-1037).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>
#include "csu_host.h"

/* The whole file in memory, read token by token with meshGetWord's
   semantics (mesh.c:10-42): a token ends at the first ' ', '\r' or '\n',
   which is consumed, so consecutive delimiters give empty tokens; returns
   the delimiter, -1 at the end of the file. */
typedef struct { const char* p; const char* end; } rbuf;

static int get_word(rbuf* f, char* s, int cap)
{
    int n = 0, c = -1;
    while (f->p < f->end) {
        c = (unsigned char)*f->p++;
        if (c == ' ' || c == '\r' || c == '\n') break;
        if (n < cap - 1) s[n++] = (char)c;
        c = -1;
    }
    s[n] = '\0';
    return c;
}

static char* slurp(const char* path, size_t* len)
{
    FILE* f = fopen(path, "rb");
    if (!f) return NULL;
    size_t cap = 1 << 20, n = 0;
    char* b = malloc(cap);
    if (!b) { fclose(f); return NULL; }
    for (;;) {
        if (n == cap) {
            char* nb = realloc(b, cap * 2);
            if (!nb) { free(b); fclose(f); return NULL; }
            b = nb;
            cap *= 2;
        }
        size_t r = fread(b + n, 1, cap - n, f);
        n += r;
        if (r == 0) break;
    }
    fclose(f);
    *len = n;
    return b;
}

/* Stable LSD radix sort of idx[0..n) by key[idx[.]] (8-bit digits, only as
   many as the largest key needs): replaces the comparison sorts of the
   face enumeration, whose keys make every order total. */
static int radix_sort(const uint64_t* key, uint32_t* idx, size_t n)
{
    uint64_t mx = 0;
    for (size_t i = 0; i < n; i++) if (key[i] > mx) mx = key[i];
    uint32_t* tmp = malloc(sizeof(uint32_t) * (n ? n : 1));
    if (!tmp) return -1;
    for (int shift = 0; shift < 64 && (mx >> shift) != 0; shift += 8) {
        size_t cnt[257] = {0};
        for (size_t i = 0; i < n; i++) cnt[((key[idx[i]] >> shift) & 255) + 1]++;
        for (int b = 0; b < 256; b++) cnt[b + 1] += cnt[b];
        for (size_t i = 0; i < n; i++) tmp[cnt[(key[idx[i]] >> shift) & 255]++] = idx[i];
        memcpy(idx, tmp, sizeof(uint32_t) * n);
    }
    free(tmp);
    return 0;
}

static double tri_area(const double* xy, int p0, int p1, int p2)
{
    /* reference meshCalcOmegaTri, mesh.c:388-401 (same operation order) */
    double x1 = xy[2 * p0], x2 = xy[2 * p1], x3 = xy[2 * p2];
    double y1 = xy[2 * p0 + 1], y2 = xy[2 * p1 + 1], y3 = xy[2 * p2 + 1];
    return 0.5 * ((x1 - x2) * (y1 + y2) + (x2 - x3) * (y2 + y3) + (x3 - x1) * (y3 + y1));
}

static double cell_volume(const double* xy, const int32_t* v, int axi)
{
    int np = v[3] >= 0 ? 4 : 3;
    double a = tri_area(xy, v[0], v[1], v[2]);
    if (np == 4) a += tri_area(xy, v[2], v[3], v[0]);
    if (axi == 1) {
        double x = 0.0, y = 0.0;
        for (int k = 0; k < np; k++) { x += xy[2 * v[k]]; y += xy[2 * v[k] + 1]; }
        y /= np;
        a *= y;
    }
    return a;
}

typedef struct { int64_t key; int64_t seq; int32_t cell, p0, p1; } dedge;
typedef struct { int64_t bucket; int64_t seq; int32_t e0, e1, p0, p1; } rface;

static int64_t hash_bucket(int64_t size, int p0, int p1)
{
    /* reference meshHashFunc, mesh.c:1083-1090 */
    double c = (sqrt(5.0) - 1.0) / 2.0;
    double aux = (double)(p0 + p1) * c;
    aux = aux - floor(aux);
    return (int64_t)floor((double)size * aux);
}

static int err(const char* msg, const char* path)
{
    fprintf(stderr, "csu mesh: %s (%s)\n", msg, path);
    return CSU_ERR_IO;
}

static int nomem(const char* path)
{
    fprintf(stderr, "csu mesh: out of host memory (%s)\n", path);
    return CSU_ERR_NOMEM;
}

int csu_mesh_read_su2(const char* path, int32_t axi, csu_case* c)
{
    size_t flen = 0;
    char* fbuf = slurp(path, &flen);
    if (!fbuf) return err("cannot open or read into memory", path);
    rbuf rb = {fbuf, fbuf + flen}, *f = &rb;
    char s[128];
    int nelem = -1, npoin = -1, nmark = -1;

    get_word(f, s, sizeof s);
    if (strcmp(s, "NDIME=") == 0) get_word(f, s, sizeof s);
    get_word(f, s, sizeof s);
    if (strcmp(s, "NELEM=") == 0) { get_word(f, s, sizeof s); nelem = atoi(s); }
    if (nelem <= 0) { free(fbuf); return err("missing NELEM=", path); }

    int32_t* cv = malloc(sizeof(int32_t) * 4 * (size_t)nelem);
    if (!cv) { free(fbuf); return nomem(path); }
    for (int i = 0; i < nelem; i++) {
        get_word(f, s, sizeof s);
        int type = atoi(s);
        int np = type == 5 ? 3 : type == 9 ? 4 : 0;
        if (!np) { free(cv); free(fbuf); return err("element type is not 5 or 9", path); }
        for (int k = 0; k < 4; k++) cv[4 * i + k] = -1;
        for (int k = 0; k < np; k++) {
            get_word(f, s, sizeof s);
            cv[4 * i + k] = atoi(s);
            /* -1 marks the missing 4th vertex of a triangle: no element may
               name a negative vertex (a quad would turn into a triangle) */
            if (cv[4 * i + k] < 0) { free(cv); free(fbuf); return err("negative vertex index", path); }
        }
        get_word(f, s, sizeof s);                       /* element index */
    }
    get_word(f, s, sizeof s);
    if (strcmp(s, "NPOIN=") == 0) { get_word(f, s, sizeof s); npoin = atoi(s); }
    if (npoin <= 0) { free(cv); free(fbuf); return err("missing NPOIN=", path); }
    double* xy = malloc(sizeof(double) * 2 * (size_t)npoin);
    if (!xy) { free(cv); free(fbuf); return nomem(path); }
    for (int i = 0; i < npoin; i++) {
        get_word(f, s, sizeof s); xy[2 * i] = strtod(s, NULL);
        get_word(f, s, sizeof s); xy[2 * i + 1] = strtod(s, NULL);
        get_word(f, s, sizeof s);                       /* point index */
    }
    get_word(f, s, sizeof s);
    if (strcmp(s, "NMARK=") == 0) { get_word(f, s, sizeof s); nmark = atoi(s); }
    if (nmark < 0) { free(cv); free(xy); free(fbuf); return err("missing NMARK=", path); }

    char (*mname)[64] = calloc((size_t)(nmark > 0 ? nmark : 1), sizeof *mname);
    int* mcount = calloc((size_t)(nmark > 0 ? nmark : 1), sizeof(int));
    int nb = 0, bcap = 1024;
    int32_t* bp = malloc(sizeof(int32_t) * 2 * (size_t)bcap);
    int32_t* bm = malloc(sizeof(int32_t) * (size_t)bcap);
    if (!mname || !mcount || !bp || !bm) {
        free(cv); free(xy); free(fbuf); free(mname); free(mcount); free(bp); free(bm);
        return nomem(path);
    }
    for (int m = 0; m < nmark; m++) {
        get_word(f, s, sizeof s);                       /* MARKER_TAG= */
        get_word(f, mname[m], sizeof mname[m]);
        get_word(f, s, sizeof s);
        int cnt = 0;
        if (strcmp(s, "MARKER_ELEMS=") == 0) { get_word(f, s, sizeof s); cnt = atoi(s); }
        mcount[m] = cnt;
        for (int j = 0; j < cnt; j++) {
            if (nb == bcap) {
                bcap *= 2;
                int32_t* nbp = realloc(bp, sizeof(int32_t) * 2 * (size_t)bcap);
                if (nbp) bp = nbp;
                int32_t* nbm = nbp ? realloc(bm, sizeof(int32_t) * (size_t)bcap) : NULL;
                if (nbm) bm = nbm;
                if (!nbp || !nbm) {
                    free(cv); free(xy); free(fbuf); free(mname); free(mcount); free(bp); free(bm);
                    return nomem(path);
                }
            }
            get_word(f, s, sizeof s);                   /* line type (3) */
            get_word(f, s, sizeof s); bp[2 * nb] = atoi(s);
            get_word(f, s, sizeof s); bp[2 * nb + 1] = atoi(s);
            get_word(f, s, sizeof s);                   /* trailing token */
            bm[nb] = m;
            nb++;
        }
    }
    free(fbuf);
    free(mcount);

    for (int i = 0; i < nelem; i++)
        for (int k = 0; k < 4; k++)
            if (cv[4 * i + k] >= npoin) {
                free(cv); free(xy); free(bp); free(bm); free(mname);
                return err("vertex index out of range", path);
            }
    for (int g = 0; g < nb; g++)
        if (bp[2 * g] < 0 || bp[2 * g] >= npoin || bp[2 * g + 1] < 0 || bp[2 * g + 1] >= npoin) {
            free(cv); free(xy); free(bp); free(bm); free(mname);
            return err("boundary point index out of range", path);
        }

    /* orientation fix, mesh.c:222-242 */
    for (int i = 0; i < nelem; i++) {
        int32_t* v = cv + 4 * i;
        if (cell_volume(xy, v, axi) < 0) {
            int np = v[3] >= 0 ? 4 : 3, tmp[4];
            for (int k = 0; k < np; k++) tmp[k] = v[np - 1 - k];
            for (int k = 0; k < np; k++) v[k] = tmp[k];
        }
    }

    /* directed edges in insertion order, paired on the undirected key */
    size_t nde = 0;
    dedge* de = malloc(sizeof(dedge) * 4 * (size_t)nelem);
    if (!de) { free(cv); free(xy); free(bp); free(bm); free(mname); return nomem(path); }
    for (int i = 0; i < nelem; i++) {
        int np = cv[4 * i + 3] >= 0 ? 4 : 3;
        for (int k = 0; k < np; k++) {
            int p0 = cv[4 * i + k], p1 = cv[4 * i + (k + 1) % np];
            int lo = p0 < p1 ? p0 : p1, hi = p0 < p1 ? p1 : p0;
            de[nde].key = (int64_t)lo * (int64_t)npoin + hi;
            de[nde].seq = (int64_t)nde;
            de[nde].cell = i; de[nde].p0 = p0; de[nde].p1 = p1;
            nde++;
        }
    }
    {
        /* sorted on (key, seq): de is built in seq order, the sort is stable */
        uint64_t* k = malloc(sizeof(uint64_t) * (nde ? nde : 1));
        uint32_t* ix = malloc(sizeof(uint32_t) * (nde ? nde : 1));
        dedge* ds = malloc(sizeof(dedge) * (nde ? nde : 1));
        if (!k || !ix || !ds) {
            free(k); free(ix); free(ds); free(de); free(cv); free(xy); free(bp); free(bm); free(mname);
            return nomem(path);
        }
        for (size_t i = 0; i < nde; i++) { k[i] = (uint64_t)de[i].key; ix[i] = (uint32_t)i; }
        if (radix_sort(k, ix, nde)) {
            free(k); free(ix); free(ds); free(de); free(cv); free(xy); free(bp); free(bm); free(mname);
            return nomem(path);
        }
        for (size_t i = 0; i < nde; i++) ds[i] = de[ix[i]];
        free(de); free(k); free(ix);
        de = ds;
    }
    rface* rf = malloc(sizeof(rface) * (nde / 2 + 1));
    if (!rf) { free(de); free(cv); free(xy); free(bp); free(bm); free(mname); return nomem(path); }
    size_t nf = 0;
    int64_t hsize = (int64_t)nelem + npoin;
    for (size_t a = 0; a < nde;) {
        size_t b = a + 1;
        while (b < nde && de[b].key == de[a].key) b++;
        if (b - a == 2) {
            if (!(de[a].p0 == de[a + 1].p1 && de[a].p1 == de[a + 1].p0)) {
                free(cv); free(xy); free(bp); free(bm); free(mname); free(de); free(rf);
                return err("cells sharing an edge with the same orientation", path);
            }
            rf[nf].bucket = hash_bucket(hsize, de[a].p0, de[a].p1);
            rf[nf].seq = de[a].seq;
            rf[nf].e0 = de[a].cell; rf[nf].e1 = de[a + 1].cell;
            rf[nf].p0 = de[a].p0; rf[nf].p1 = de[a].p1;
            nf++;
        } else if (b - a > 2) {
            free(cv); free(xy); free(bp); free(bm); free(mname); free(de); free(rf);
            return err("edge shared by more than two cells", path);
        }
        a = b;
    }
    free(de);
    {
        /* sorted on (bucket, seq): both below 2^31, one 62-bit key */
        uint64_t* k = malloc(sizeof(uint64_t) * (nf ? nf : 1));
        uint32_t* ix = malloc(sizeof(uint32_t) * (nf ? nf : 1));
        rface* rs = malloc(sizeof(rface) * (nf + 1));
        if (!k || !ix || !rs) {
            free(k); free(ix); free(rs); free(rf); free(cv); free(xy); free(bp); free(bm); free(mname);
            return nomem(path);
        }
        for (size_t i = 0; i < nf; i++) {
            k[i] = ((uint64_t)rf[i].bucket << 31) | (uint64_t)rf[i].seq;
            ix[i] = (uint32_t)i;
        }
        if (radix_sort(k, ix, nf)) {
            free(k); free(ix); free(rs); free(rf); free(cv); free(xy); free(bp); free(bm); free(mname);
            return nomem(path);
        }
        for (size_t i = 0; i < nf; i++) rs[i] = rf[ix[i]];
        free(rf); free(k); free(ix);
        rf = rs;
    }

    int32_t* face = malloc(sizeof(int32_t) * 4 * (nf ? nf : 1));
    int32_t* nbr = malloc(sizeof(int32_t) * 4 * (size_t)nelem);
    int32_t* cface = malloc(sizeof(int32_t) * 4 * (size_t)nelem);
    int* nn = calloc((size_t)nelem, sizeof(int));
    if (!face || !nbr || !cface || !nn) {
        free(face); free(nbr); free(cface); free(nn); free(rf); free(cv); free(xy); free(bp); free(bm); free(mname);
        return nomem(path);
    }
    for (int i = 0; i < 4 * nelem; i++) { nbr[i] = INT32_MAX; cface[i] = 0; }
    for (size_t j = 0; j < nf; j++) {
        face[4 * j + 0] = rf[j].e0; face[4 * j + 1] = rf[j].e1;
        face[4 * j + 2] = rf[j].p0; face[4 * j + 3] = rf[j].p1;
        int e0 = rf[j].e0, e1 = rf[j].e1;
        if (nn[e0] >= 4 || nn[e1] >= 4) {
            free(cv); free(xy); free(bp); free(bm); free(mname); free(rf);
            free(face); free(nbr); free(cface); free(nn);
            return err("cell with more than four neighbours", path);
        }
        cface[4 * e0 + nn[e0]] = (int32_t)(j + 1);
        cface[4 * e1 + nn[e1]] = -(int32_t)(j + 1);
        nbr[4 * e0 + nn[e0]++] = e1;
        nbr[4 * e1 + nn[e1]++] = e0;
    }
    free(rf);

    /* node -> cells incidence (ascending cell ids) */
    int* ninc = calloc((size_t)npoin + 1, sizeof(int));
    if (!ninc) {
        free(face); free(nbr); free(cface); free(nn); free(cv); free(xy); free(bp); free(bm); free(mname);
        return nomem(path);
    }
    for (int i = 0; i < nelem; i++)
        for (int k = 0; k < 4; k++) if (cv[4 * i + k] >= 0) ninc[cv[4 * i + k] + 1]++;
    for (int p = 0; p < npoin; p++) ninc[p + 1] += ninc[p];
    int* inc = malloc(sizeof(int) * (size_t)(ninc[npoin] > 0 ? ninc[npoin] : 1));
    int* fill = calloc((size_t)npoin, sizeof(int));
    if (!inc || !fill) {
        free(inc); free(fill); free(ninc); free(face); free(nbr); free(cface); free(nn); free(cv); free(xy);
        free(bp); free(bm); free(mname);
        return nomem(path);
    }
    for (int i = 0; i < nelem; i++)
        for (int k = 0; k < 4; k++) {
            int p = cv[4 * i + k];
            if (p >= 0) inc[ninc[p] + fill[p]++] = i;
        }
    free(fill);

    int32_t* bface = malloc(sizeof(int32_t) * 3 * (size_t)(nb ? nb : 1));
    if (!bface) {
        free(inc); free(ninc); free(face); free(nbr); free(cface); free(nn); free(cv); free(xy); free(bp); free(bm);
        free(mname);
        return nomem(path);
    }
    for (int g = 0; g < nb; g++) {
        int p0 = bp[2 * g], p1 = bp[2 * g + 1], e = -1;
        for (int t = ninc[p0]; t < ninc[p0 + 1] && e < 0; t++) {
            int ci = inc[t], links = 0;
            for (int k = 0; k < 4; k++) {
                int v = cv[4 * ci + k];
                if (v < 0) continue;
                links += (v == p0) + (v == p1);
            }
            if (links == 2) e = ci;
        }
        if (e < 0 || nn[e] >= 4) {
            free(cv); free(xy); free(bp); free(bm); free(mname); free(face);
            free(nbr); free(cface); free(nn); free(ninc); free(inc); free(bface);
            return err(e < 0 ? "boundary face without a cell" : "cell with more than four neighbours", path);
        }
        nbr[4 * e + nn[e]++] = -(1 + g);
        /* orientation: (p0,p1) must be a CCW edge of e (mesh.c:948-963, 987-1009) */
        int np = cv[4 * e + 3] >= 0 ? 4 : 3, ori = -1;
        for (int k = 0; k < np; k++)
            if (cv[4 * e + k] == p0 && cv[4 * e + (k + 1) % np] == p1) ori = 1;
        if (ori < 0) { int t = p0; p0 = p1; p1 = t; }
        bface[3 * g + 0] = e; bface[3 * g + 1] = p0; bface[3 * g + 2] = p1;
    }
    free(ninc); free(inc); free(nn); free(bp);

    c->mesh.ncell = nelem;
    c->mesh.nnode = npoin;
    c->mesh.nface = (int32_t)nf;
    c->mesh.nbface = nb;
    c->mesh.nmark = nmark;
    c->mesh.axi = axi;
    c->node_xy = xy; c->cell_vert = cv; c->face = face; c->bface = bface;
    c->bface_mark = bm; c->cell_nbr = nbr; c->cell_face = cface;
    c->mark_name = mname;
    c->mark_type = calloc((size_t)(nmark > 0 ? nmark : 1), sizeof(int32_t));
    if (!c->mark_type) return nomem(path);  /* the arrays above belong to c: csu_case_free releases them */
    c->mesh.node_xy = xy; c->mesh.cell_vert = cv; c->mesh.face = face;
    c->mesh.bface = bface; c->mesh.bface_mark = bm; c->mesh.mark_type = c->mark_type;
    c->mesh.cell_nbr = nbr; c->mesh.cell_face = cface; c->mesh.wall_dist = NULL;
    return CSU_OK;
}

int csu_wall_distance(csu_case* c)
{
    /* reference saCalcD, sa.c:703-763: min over wall/wallT faces of the
       point-to-segment distance from each cell's vertex average */
    const csu_mesh_desc* m = &c->mesh;
    double* d = malloc(sizeof(double) * (size_t)m->ncell);
    if (!d) return CSU_ERR_NOMEM;
    for (int i = 0; i < m->ncell; i++) {
        const int32_t* v = m->cell_vert + 4 * i;
        int np = v[3] >= 0 ? 4 : 3;
        double xc = 0.0, yc = 0.0;
        for (int k = 0; k < np; k++) { xc += m->node_xy[2 * v[k]]; yc += m->node_xy[2 * v[k] + 1]; }
        xc /= np; yc /= np;
        double best = 1e6;
        for (int g = 0; g < m->nbface; g++) {
            int t = m->mark_type[m->bface_mark[g]];
            if (t != CSU_BC_WALL && t != CSU_BC_WALLT) continue;
            double p0x = m->node_xy[2 * m->bface[3 * g + 1]], p0y = m->node_xy[2 * m->bface[3 * g + 1] + 1];
            double p1x = m->node_xy[2 * m->bface[3 * g + 2]], p1y = m->node_xy[2 * m->bface[3 * g + 2] + 1];
            double num = (p1x - p0x) * (xc - p0x) + (p1y - p0y) * (yc - p0y);
            double den = (p1x - p0x) * (p1x - p0x) + (p1y - p0y) * (p1y - p0y);
            double tt = num / den;
            if (tt > 1.) tt = 1.; else if (tt < 0.) tt = 0.;
            double dx = (p1x - p0x) * tt + p0x - xc;
            double dy = (p1y - p0y) * tt + p0y - yc;
            double dd = sqrt(dx * dx + dy * dy);
            if (best > dd) best = dd;
        }
        d[i] = best;
    }
    free(c->wall_dist);
    c->wall_dist = d;
    c->mesh.wall_dist = d;
    return CSU_OK;
}
