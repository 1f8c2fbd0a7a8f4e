/*
 * cw_tree.c -- CPU restatement of ClustalW 2.0.11's neighbour-joining guide tree
 * and its PHYLIP (.dnd) writer.  TEST INFRASTRUCTURE ONLY (the checker that
 * compares a DistMatrix with the reference's own .dnd fixtures).
 *
 * Restates (paths relative to CPU/clustalw-2.0.11/src/):
 *   tree/NJTree.cpp:161-751              NJTree::generateTree (incl. the
 *                                        incremental row/column sums and the
 *                                        linked list of live nodes, whose
 *                                        summation order fixes the doubles)
 *   tree/ClusterTreeOutput.cpp:45-220    printPhylipTree / twoWaySplit
 *   tree/UnRootedClusterTree.cpp:170-257 treeFromDistMatrix (the 2-sequence case)
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    char *buf;
    size_t len, cap;
} sbuf;

static void sb_puts(sbuf *s, const char *t)
{
    size_t n = strlen(t);
    if (s->len + n + 1 > s->cap) {
        size_t nc = (s->cap + n + 1) * 2;
        s->buf = (char *)realloc(s->buf, nc);
        s->cap = nc;
    }
    memcpy(s->buf + s->len, t, n + 1);
    s->len += n;
}

static void sb_fixed5(sbuf *s, double v)
{
    char t[64];
    snprintf(t, sizeof t, "%.5f", v);     /* iostream `fixed << setprecision(5)` */
    sb_puts(s, t);
}

typedef struct {
    int nseq;
    double *d;          /* (n+1) x (n+1) symmetric, 1-based (SymMatrix semantics) */
    int *desc;          /* treeDesc (n+1) x (n+1) */
    double *left, *right;
    const char **names;
} tree_t;

#define D(t, i, j) ((t)->d[(size_t)(i) * ((t)->nseq + 1) + (j)])
#define DESC(t, i, j) ((t)->desc[(size_t)(i) * ((t)->nseq + 1) + (j)])

/* SymMatrix::SetAt on the symmetric store */
static void set_at(tree_t *t, int i, int j, double v)
{
    D(t, i, j) = v;
    D(t, j, i) = v;
}

/* NJTree::generateTree (NJTree.cpp:161-751), firstSeq = 1 */
static void nj(tree_t *t)
{
    const int last = t->nseq;
    int i, j, k, nc, mini, minj, ii, jj, l[4], nude;
    double fnseqs, fnseqs2, sumd, diq, djq, dij, dio, djo, da, tmin, total, dmin;
    double bi, bj, b1, b2, b3, branch[4];
    int typei, typej;
    double *sumCols = (double *)calloc(last + 1, sizeof(double));
    double *sumRows = (double *)calloc(last + 1, sizeof(double));
    double *join = (double *)calloc(last + 1, sizeof(double));
    double *av = (double *)calloc(last + 1, sizeof(double));
    int *tkill = (int *)calloc(last + 1, sizeof(int));
    /* live-node chain: node 0 is the header */
    int *nxt = (int *)malloc((last + 2) * sizeof(int));
    int *prv = (int *)malloc((last + 2) * sizeof(int));
    const int NIL = -1;

    fnseqs = (double)last;
    for (i = 0; i <= last; i++) {
        nxt[i] = (i < last) ? i + 1 : NIL;
        prv[i] = (i > 0) ? i - 1 : NIL;
    }
    for (i = 1; i <= last; i++) {
        D(t, i, i) = av[i] = 0.0;
        tkill[i] = 0;
    }
    /* initial sums (:244-274) */
    sumd = 0.0;
    for (j = nxt[0]; j != NIL; j = nxt[j]) {
        double tmp = 0.0;
        int p = nxt[0];
        for (; p != NIL && p < j; p = nxt[p]) tmp = tmp + D(t, p, j);
        sumRows[j] = tmp;
        tmp = 0.0;
        if (p != NIL && p == j) p = nxt[p];
        for (; p != NIL; p = nxt[p]) tmp += D(t, j, p);
        sumCols[j] = tmp;
    }

    for (nc = 1; nc <= last - 3; ++nc) {
        sumd = 0.0;
        for (j = nxt[0]; j != NIL; j = nxt[j]) sumd += sumCols[j];
        fnseqs2 = fnseqs - 2.0;
        tmin = 99999.0 * 2.0 * fnseqs2;
        mini = minj = 0;
        jj = (nxt[0] != NIL && nxt[0] == 1) ? nxt[nxt[0]] : nxt[0];
        for (; jj != NIL; jj = nxt[jj]) {
            for (ii = nxt[0]; ii < jj; ii = nxt[ii]) {
                diq = sumCols[ii] + sumRows[ii];
                djq = sumCols[jj] + sumRows[jj];
                dij = D(t, ii, jj);
                total = fnseqs2 * dij + 2.0 * sumd - diq - djq;
                if (total < tmin) {
                    tmin = total;
                    mini = ii;
                    minj = jj;
                }
            }
        }
        dio = sumCols[mini] + sumRows[mini];
        djo = sumCols[minj] + sumRows[minj];
        dmin = D(t, mini, minj);
        dio = (dio - dmin) / fnseqs2;
        djo = (djo - dmin) / fnseqs2;
        bi = (dmin + dio - djo) * 0.5;
        bj = dmin - bi;
        bi = bi - av[mini];
        bj = bj - av[minj];
        typei = (av[mini] > 0.0) ? 0 : 1;
        typej = (av[minj] > 0.0) ? 0 : 1;
        if (fabs(bi) < 0.0001) bi = 0.0;
        if (fabs(bj) < 0.0001) bj = 0.0;
        t->left[nc] = bi;
        t->right[nc] = bj;
        for (i = 1; i <= last; i++) DESC(t, nc, i) = 0;
        if (typei == 0) {
            for (i = nc - 1; i >= 1; i--)
                if (DESC(t, i, mini) == 1) {
                    for (j = 1; j <= last; j++)
                        if (DESC(t, i, j) == 1) DESC(t, nc, j) = 1;
                    break;
                }
        } else {
            DESC(t, nc, mini) = 1;
        }
        if (typej == 0) {
            for (i = nc - 1; i >= 1; i--)
                if (DESC(t, i, minj) == 1) {
                    for (j = 1; j <= last; j++)
                        if (DESC(t, i, j) == 1) DESC(t, nc, j) = 1;
                    break;
                }
        } else {
            DESC(t, nc, minj) = 1;
        }
        if (dmin <= 0.0) dmin = 0.000001;
        av[mini] = dmin * 0.5;
        fnseqs = fnseqs - 1.0;
        tkill[minj] = 1;
        /* unlink minj */
        nxt[prv[minj]] = nxt[minj];
        if (nxt[minj] != NIL) prv[nxt[minj]] = prv[minj];
        prv[minj] = nxt[minj] = NIL;
        for (j = nxt[0]; j != NIL; j = nxt[j]) {
            double tdi = 0.0, tdj = 0.0;
            if (j < minj) {
                tdj = D(t, j, minj);
                sumCols[j] -= tdj;
            } else if (j > minj) {
                tdj = D(t, minj, j);
                sumRows[j] -= tdj;
            }
            if (j < mini) {
                tdi = D(t, j, mini);
                sumCols[j] -= tdi;
            } else if (j > mini) {
                tdi = D(t, mini, j);
                sumRows[j] -= tdi;
            }
            join[j] = (tdj + tdi) * 0.5;
        }
        for (j = nxt[0]; j != NIL; j = nxt[j]) {
            if (j < mini) {
                set_at(t, j, mini, join[j]);
                sumCols[j] += join[j];
            } else if (j > mini) {
                set_at(t, mini, j, join[j]);
                sumRows[j] += join[j];
            }
        }
        sumCols[mini] = sumRows[mini] = 0.0;
        da = 0.0;
        j = nxt[0];
        for (; j < mini; j = nxt[j]) da = da + join[j];
        sumRows[mini] = da;
        if (j != NIL && j == mini) j = nxt[j];
        da = 0.0;
        for (; j != NIL; j = nxt[j]) da = da + join[j];
        sumCols[mini] = da;
        sumCols[minj] = sumRows[minj] = 0.0;
        for (j = 1; j <= last; ++j) {
            set_at(t, minj, j, 0.0);
            set_at(t, j, minj, 0.0);
            join[j] = 0.0;
        }
    }

    /* last cycle: trichotomy (:615-741) */
    nude = 1;
    for (i = nxt[0]; i != NIL; i = nxt[i]) l[nude++] = i;
    b1 = (D(t, l[1], l[2]) + D(t, l[1], l[3]) - D(t, l[2], l[3])) * 0.5;
    b2 = D(t, l[1], l[2]) - b1;
    b3 = D(t, l[1], l[3]) - b1;
    branch[1] = b1 - av[l[1]];
    branch[2] = b2 - av[l[2]];
    branch[3] = b3 - av[l[3]];
    for (k = 1; k <= 3; k++)
        if (fabs(branch[k]) < 0.0001) branch[k] = 0.0;
    t->left[last - 2] = branch[1];
    t->left[last - 1] = branch[2];
    t->left[last] = branch[3];
    for (i = 1; i <= last; i++) DESC(t, last - 2, i) = 0;
    for (i = 1; i <= 3; ++i) {
        if (av[l[i]] > 0.0) {
            for (k = last - 3; k >= 1; k--)
                if (DESC(t, k, l[i]) == 1) {
                    for (j = 1; j <= last; j++)
                        if (DESC(t, k, j) == 1) DESC(t, last - 2, j) = i;
                    break;
                }
        } else {
            DESC(t, last - 2, l[i]) = i;
        }
    }
    free(sumCols); free(sumRows); free(join); free(av); free(tkill); free(nxt); free(prv);
}

/* ClusterTreeOutput::twoWaySplit (:101-220), bootstrap labels off */
static int two_way_split(tree_t *t, sbuf *s, int startRow, int flag)
{
    const int last = t->nseq;
    int row, newRow = 0, col, testCol = 0, single;
    if (startRow != last - 2) sb_puts(s, "(\n");
    for (col = 1; col <= last; col++)
        if (DESC(t, startRow, col) == flag) { testCol = col; break; }
    single = 1;
    for (row = startRow - 1; row >= 1; row--)
        if (DESC(t, row, testCol) == 1) { single = 0; newRow = row; break; }
    if (single) {
        DESC(t, startRow, testCol) = 0;
        sb_puts(s, t->names[testCol]);
        if (startRow == last - 2) return 0;
        sb_puts(s, ":");
        sb_fixed5(s, t->left[startRow]);
        sb_puts(s, ",\n");
    } else {
        for (col = 1; col <= last; col++)
            if (DESC(t, startRow, col) == 1 && DESC(t, newRow, col) == 1) DESC(t, startRow, col) = 0;
        two_way_split(t, s, newRow, 1);
        if (startRow == last - 2) return newRow;
        sb_puts(s, ":");
        sb_fixed5(s, t->left[startRow]);
        sb_puts(s, ",\n");
    }
    for (col = 1; col <= last; col++)
        if (DESC(t, startRow, col) == flag) { testCol = col; break; }
    single = 1;
    newRow = 0;
    for (row = startRow - 1; row >= 1; row--)
        if (DESC(t, row, testCol) == 1) { single = 0; newRow = row; break; }
    if (single) {
        DESC(t, startRow, testCol) = 0;
        sb_puts(s, t->names[testCol]);
        sb_puts(s, ":");
        sb_fixed5(s, t->right[startRow]);
        sb_puts(s, ")\n");
    } else {
        for (col = 1; col <= last; col++)
            if (DESC(t, startRow, col) == 1 && DESC(t, newRow, col) == 1) DESC(t, startRow, col) = 0;
        two_way_split(t, s, newRow, 1);
        sb_puts(s, ":");
        sb_fixed5(s, t->right[startRow]);
        sb_puts(s, ")\n");
    }
    return startRow;
}

/* The guide tree text of UnRootedClusterTree::treeFromDistMatrix for sequences
 * 1..n.  dist is (n+1)x(n+1) row-major (entry [i][j] = distance of sequences
 * i and j, 1-based; it is modified).  names[1..n].  Returns a malloc'd string. */
char *cwo_nj_dnd(double *dist, int n, const char **names)
{
    sbuf s = {0, 0, 0};
    tree_t t;
    sb_puts(&s, "");
    if (n < 2) return s.buf;
    if (n == 2) {
        /* :225-233: default (not fixed) 5-significant-digit output */
        char tmp[128];
        double d2 = dist[1 * (n + 1) + 2] / 2.0;
        sb_puts(&s, "(");
        sb_puts(&s, names[1]);
        snprintf(tmp, sizeof tmp, ":%.5g,", (double)(float)d2);
        sb_puts(&s, tmp);
        sb_puts(&s, names[2]);
        snprintf(tmp, sizeof tmp, ":%.5g);\n", (double)(float)d2);
        sb_puts(&s, tmp);
        return s.buf;
    }
    t.nseq = n;
    t.d = dist;
    t.names = names;
    t.desc = (int *)calloc((size_t)(n + 1) * (n + 1), sizeof(int));
    t.left = (double *)calloc(n + 2, sizeof(double));
    t.right = (double *)calloc(n + 2, sizeof(double));
    nj(&t);
    /* printPhylipTree (:45-99) */
    sb_puts(&s, "(\n");
    two_way_split(&t, &s, n - 2, 1);
    sb_puts(&s, ":");
    sb_fixed5(&s, t.left[n - 2]);
    sb_puts(&s, ",\n");
    two_way_split(&t, &s, n - 2, 2);
    sb_puts(&s, ":");
    sb_fixed5(&s, t.left[n - 1]);
    sb_puts(&s, ",\n");
    two_way_split(&t, &s, n - 2, 3);
    sb_puts(&s, ":");
    sb_fixed5(&s, t.left[n]);
    sb_puts(&s, ");\n");
    free(t.desc);
    free(t.left);
    free(t.right);
    return s.buf;
}

void cwo_free(void *p) { free(p); }
