/*
 * cw_guide.c -- CPU restatement of how ClustalW 2.0.11 turns a guide tree
 * (.dnd text) into the inputs of the progressive alignment: the sequence
 * weights, the similarity matrix and the list of alignment steps.
 * TEST INFRASTRUCTURE ONLY (checker for the progressive stage).
 *
 * Restates (paths relative to CPU/clustalw-2.0.11/src/):
 *   tree/Tree.cpp:104-265   readTree (leaf/node parser, trifurcation insert,
 *                           re-rooting, leaf order by name)
 *   tree/Tree.cpp:470-587   createTree / createNode / insertNode
 *   tree/Tree.cpp:701-888   reRoot / insertRoot (incl. removal of the old root)
 *   tree/Tree.cpp:896-1058  calcRootMean / calcMean
 *   tree/Tree.cpp:1064-1097 orderNodes / calcWeight
 *   tree/Tree.cpp:22-100    calcSeqWeights
 *   tree/Tree.cpp:327-469   calcSimilarities
 *   tree/Tree.cpp:273-322   createSets, :1120-1226 groupSeqs / markGroup1 / markGroup2
 *   tree/TreeInterface.cpp:209-276 the order of the calls above
 *
 * All distances are single-precision floats, as in the reference's TreeNode;
 * the library must be built without fast-math and with FP contraction off so
 * every float expression rounds as written.
 */
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cw_oracle.h"

#define NODE 0
#define LEAF 1
#define DIR_LEFT 1
#define DIR_RIGHT 2
#define MAXNAMES 150          /* general/clustalw.h:164 */
#define MAXERRS 10            /* tree/Tree.h */
#define INT_SCALE_FACTOR 1000 /* general/clustalw.h:226 (OS_UNIX) */

typedef struct tnode {
    struct tnode *parent, *left, *right;
    int leaf;
    int order;
    float dist;
    char name[MAXNAMES + 1];
} tnode;

typedef struct {
    const char *txt;
    size_t pos, len;
    int ch;                   /* charFromFile */
    tnode **nptr, **ptrs, **lptr, **olptr;
    int nnodes, ntotal, numSeq;
    int rooted;
    int distance_tree;
    tnode *seqTree, *root;
    /* every node ever allocated, freed at the end */
    tnode **all;
    int nall, capall;
} tree_rd;

static int rd_get(tree_rd *t)
{
    if (t->pos >= t->len) { t->pos++; return -1; }
    return (unsigned char)t->txt[t->pos++];
}

/* Tree::skipSpace: get until a non-space, put it back */
static void rd_skip_space(tree_rd *t)
{
    int c;
    do {
        c = rd_get(t);
    } while (c >= 0 && isspace(c));
    t->pos--;
}

/* `file >> dist` for a float */
static float rd_float(tree_rd *t)
{
    const char *s = t->txt + t->pos;
    char *end;
    float v = strtof(s, &end);
    t->pos += (size_t)(end - s);
    return v;
}

/* Tree::avail */
static tnode *avail(tree_rd *t)
{
    tnode *p = (tnode *)calloc(1, sizeof(tnode));
    if (t->nall == t->capall) {
        t->capall = t->capall ? 2 * t->capall : 256;
        t->all = (tnode **)realloc(t->all, sizeof(tnode *) * t->capall);
    }
    t->all[t->nall++] = p;
    return p;
}

/* Tree::setInfo */
static void set_info(tnode *p, tnode *parent, int leaf, const char *name, float dist)
{
    p->parent = parent;
    p->leaf = leaf;
    p->dist = dist;
    p->order = 0;
    snprintf(p->name, sizeof p->name, "%s", name);
    if (p->leaf == LEAF) {
        p->left = NULL;
        p->right = NULL;
    }
}

/* Tree::createNode */
static void create_node(tree_rd *t, tnode *p, tnode *parent)
{
    p->parent = parent;
    p->left = avail(t);
    p->right = avail(t);
}

/* Tree::insertNode (the node that holds the third branch of a trifurcation) */
static tnode *insert_node(tree_rd *t, tnode *p)
{
    tnode *nn = avail(t);
    create_node(t, nn, p->parent);
    nn->left = p;
    p->parent = nn;
    set_info(nn, p->parent, NODE, "", 0.0f);
    return nn;
}

/* Tree::createTree */
static void create_tree(tree_rd *t, tnode *ptree, tnode *parent)
{
    int type, i;
    float dist;
    char name[MAXNAMES + 1];
    int nl = 0;
    rd_skip_space(t);
    t->ch = rd_get(t);
    if (t->ch == '(') {
        type = NODE;
        name[0] = 0;
        t->ptrs[t->ntotal] = t->nptr[t->nnodes] = ptree;
        t->nnodes++;
        t->ntotal++;
        create_node(t, ptree, parent);
        create_tree(t, ptree->left, ptree);
        if (t->ch == ',') {
            create_tree(t, ptree->right, ptree);
            if (t->ch == ',') {
                ptree = insert_node(t, ptree);
                t->ptrs[t->ntotal] = t->nptr[t->nnodes] = ptree;
                t->nnodes++;
                t->ntotal++;
                create_tree(t, ptree->right, ptree);
                t->rooted = 0;
            }
        }
        rd_skip_space(t);
        t->ch = rd_get(t);
    } else {
        type = LEAF;
        t->ptrs[t->ntotal++] = t->lptr[t->numSeq++] = ptree;
        name[nl++] = (char)t->ch;
        t->ch = rd_get(t);
        i = 1;
        while (t->ch != ':' && t->ch != ',' && t->ch != ')' && t->ch >= 0) {
            if (i < MAXNAMES) {
                name[nl++] = (char)t->ch;
                i++;
            }
            t->ch = rd_get(t);
        }
        name[nl] = 0;
        if (t->ch != ':') t->distance_tree = 0;
    }
    dist = 0.0f;
    if (t->ch == ':') {
        rd_skip_space(t);
        dist = rd_float(t);
        rd_skip_space(t);
        t->ch = rd_get(t);
    }
    set_info(ptree, parent, type, name, dist);
}

/* Tree::calcMean */
static float calc_mean(tree_rd *t, tnode *nptr, float *maxDist, int nSeqs)
{
    float dist, leftSum = 0.0f, rightSum = 0.0f, leftMean, rightMean, diff;
    tnode *p;
    tnode **pathToRoot = (tnode **)malloc(sizeof(tnode *) * nSeqs);
    float *distToNode = (float *)malloc(sizeof(float) * nSeqs);
    int depth = 0, i, j, n = 0, numLeft = 0, numRight = 0, direction, found;
    *maxDist = dist = 0.0f;
    p = nptr;
    while (p != NULL) {
        pathToRoot[depth] = p;
        dist += p->dist;
        distToNode[depth] = dist;
        p = p->parent;
        depth++;
    }
    for (i = 0; i < t->numSeq; i++) {
        p = t->lptr[i];
        if (p == nptr) {
            direction = DIR_RIGHT;
            dist = 0.0f;
        } else {
            direction = DIR_LEFT;
            dist = 0.0f;
            found = 0;
            n = 0;
            while (!found && p->parent != NULL) {
                for (j = 0; j < depth; j++)
                    if (p->parent == pathToRoot[j]) {
                        found = 1;
                        n = j;
                    }
                dist += p->dist;
                p = p->parent;
            }
            if (p == nptr) direction = DIR_RIGHT;
        }
        if (direction == DIR_LEFT) {
            leftSum += dist;
            leftSum += distToNode[n - 1];   /* n >= 1 whenever the leaf is LEFT */
            numLeft++;
        } else {
            rightSum += dist;
            numRight++;
        }
        if (dist > *maxDist) *maxDist = dist;
    }
    free(distToNode);
    free(pathToRoot);
    leftMean = leftSum / numLeft;
    rightMean = rightSum / numRight;
    diff = leftMean - rightMean;
    return diff;
}

/* Tree::insertRoot */
static tnode *insert_root(tree_rd *t, tnode *p, float diff)
{
    tnode *newp, *prev, *q, *tt;
    float dist, prevDist, td;
    newp = avail(t);
    tt = p->parent;
    prevDist = tt->dist;
    p->parent = newp;
    dist = p->dist;
    p->dist = diff / 2;
    if (p->dist < 0.0) p->dist = 0.0f;
    if (p->dist > dist) p->dist = dist;
    tt->dist = dist - p->dist;
    newp->left = tt;
    newp->right = p;
    newp->parent = NULL;
    newp->dist = 0.0f;
    newp->leaf = NODE;
    if (tt->left == p)
        tt->left = tt->parent;
    else
        tt->right = tt->parent;
    prev = tt;
    q = tt->parent;
    tt->parent = newp;
    while (q != NULL) {
        if (q->left == prev) {
            q->left = q->parent;
            q->parent = prev;
            td = q->dist;
            q->dist = prevDist;
            prevDist = td;
            prev = q;
            q = q->left;
        } else {
            q->right = q->parent;
            q->parent = prev;
            td = q->dist;
            q->dist = prevDist;
            prevDist = td;
            prev = q;
            q = q->right;
        }
    }
    /* remove the old root node */
    q = prev;
    if (q->left == NULL) {
        dist = q->dist;
        q = q->right;
        q->dist += dist;
        q->parent = prev->parent;
        if (prev->parent->left == prev)
            prev->parent->left = q;
        else
            prev->parent->right = q;
        prev->right = NULL;
    } else {
        dist = q->dist;
        q = q->left;
        q->dist += dist;
        q->parent = prev->parent;
        if (prev->parent->left == prev)
            prev->parent->left = q;
        else
            prev->parent->right = q;
        prev->left = NULL;
    }
    return newp;
}

/* Tree::reRoot */
static tnode *re_root(tree_rd *t, tnode *ptree, int nseqs)
{
    tnode *p, *rootPtr = ptree;
    float diff, minDiff = 0.0f, minDepth = 1.0f, maxDist = 0.0f;
    int i, first = 1;
    for (i = 0; i < t->ntotal; i++) {
        p = t->ptrs[i];
        if (p->parent == NULL) continue;
        diff = calc_mean(t, p, &maxDist, nseqs);
        if (diff == 0 || (diff > 0 && diff < 2 * p->dist)) {
            if (maxDist < minDepth || first) {
                first = 0;
                rootPtr = p;
                minDepth = maxDist;
                minDiff = diff;
            }
        }
    }
    if (rootPtr == ptree) {
        minDiff = rootPtr->left->dist + rootPtr->right->dist;
        rootPtr = rootPtr->right;
    }
    return insert_root(t, rootPtr, minDiff);
}

/* Tree::orderNodes */
static void order_nodes(tree_rd *t)
{
    for (int i = 0; i < t->numSeq; i++)
        for (tnode *p = t->lptr[i]; p != NULL; p = p->parent) p->order++;
}

/* Tree::calcWeight */
static int calc_weight(tree_rd *t, int leaf)
{
    float weight = 0.0f;
    tnode *p = t->olptr[leaf];
    while (p->parent != NULL) {
        weight += p->dist / p->order;
        p = p->parent;
    }
    weight *= 100.0;
    return (int)weight;
}

static void fold_name(const char *in, char *out, int spaces_to_underscore)
{
    int k = 0;
    for (; in[k] && k < MAXNAMES; k++) {
        char c = in[k];
        if (c > 0x40 && c < 0x5b) c = (char)(c | 0x20);
        if (spaces_to_underscore && c == ' ') c = '_';
        out[k] = c;
    }
    out[k] = 0;
}

/* Tree::markGroup1 / markGroup2 */
static void mark_group1(tree_rd *t, tnode *p, int *groups, int n)
{
    for (int i = 0; i < n; i++) groups[i] = (t->olptr[i] == p) ? 1 : 0;
}

static void mark_group2(tree_rd *t, tnode *p, int *groups, int n)
{
    for (int i = 0; i < n; i++) {
        if (t->olptr[i] == p)
            groups[i] = 2;
        else if (groups[i] != 0)
            groups[i] = 1;
    }
}

typedef struct {
    int *steps;     /* nsteps x (n+1) */
    int nsteps, cap, n;
} steps_t;

/* AlignmentSteps::saveSet(n, groups): element 0 is 0, then groups[0..n-1] */
static void save_set(steps_t *s, const int *groups)
{
    if (s->nsteps == s->cap) {
        s->cap = s->cap ? 2 * s->cap : 64;
        s->steps = (int *)realloc(s->steps, sizeof(int) * (size_t)s->cap * (s->n + 1));
    }
    int *row = s->steps + (size_t)s->nsteps * (s->n + 1);
    row[0] = 0;
    for (int i = 1; i <= s->n; i++) row[i] = groups[i - 1];
    s->nsteps++;
}

/* Tree::groupSeqs */
static void group_seqs(tree_rd *t, tnode *p, int *nextGroups, int nSeqs, steps_t *s)
{
    int i;
    int *tmp = (int *)calloc((size_t)nSeqs + 1, sizeof(int));
    if (p->left != NULL) {
        if (p->left->leaf == NODE) {
            group_seqs(t, p->left, nextGroups, nSeqs, s);
            for (i = 0; i < nSeqs; i++)
                if (nextGroups[i] != 0) tmp[i] = 1;
        } else {
            mark_group1(t, p->left, tmp, nSeqs);
        }
    }
    if (p->right != NULL) {
        if (p->right->leaf == NODE) {
            group_seqs(t, p->right, nextGroups, nSeqs, s);
            for (i = 0; i < nSeqs; i++)
                if (nextGroups[i] != 0) tmp[i] = 2;
        } else {
            mark_group2(t, p->right, tmp, nSeqs);
        }
        save_set(s, tmp);
    }
    for (i = 0; i < nSeqs; i++) nextGroups[i] = tmp[i];
    free(tmp);
}

int cwo_guide_from_dnd(const char *dnd, int n, const char **names, int *weights,
                       double *sim, int **steps_out, int *nsteps_out)
{
    tree_rd T;
    int i, j, k, status = 1;
    memset(&T, 0, sizeof T);
    *steps_out = NULL;
    *nsteps_out = 0;
    if (n < 2) return CWO_GUIDE_TOO_FEW;
    T.txt = dnd;
    T.len = strlen(dnd);
    T.rooted = 1;
    T.distance_tree = 1;
    /* readTree (Tree.cpp:104-265) with firstSeq = 0, lastSeq = n */
    rd_skip_space(&T);
    if (rd_get(&T) != '(') return CWO_GUIDE_FORMAT;
    T.pos = 0;
    T.nptr = (tnode **)calloc(3 * (size_t)(n + 1), sizeof(tnode *));
    T.ptrs = (tnode **)calloc(3 * (size_t)(n + 1), sizeof(tnode *));
    T.lptr = (tnode **)calloc((size_t)n + 1, sizeof(tnode *));
    T.olptr = (tnode **)calloc((size_t)n + 1, sizeof(tnode *));
    T.seqTree = avail(&T);
    set_info(T.seqTree, NULL, 0, "", 0.0f);
    create_tree(&T, T.seqTree, NULL);
    if (T.numSeq != n) {
        status = CWO_GUIDE_MISMATCH;
        goto done;
    }
    if (!T.distance_tree && T.rooted == 0) {
        status = CWO_GUIDE_FORMAT;
        goto done;
    }
    T.root = T.rooted ? T.seqTree : re_root(&T, T.seqTree, n + 1);
    order_nodes(&T);
    for (i = 0; i < n; i++) {
        char a[MAXNAMES + 1], b[MAXNAMES + 1];
        int found = 0;
        fold_name(names[i], b, 1);
        for (j = 0; j < T.numSeq; j++) {
            fold_name(T.lptr[j]->name, a, 0);
            if (strcmp(a, b) == 0) {
                T.olptr[i] = T.lptr[j];
                found = 1;
            }
        }
        if (!found) {
            status = CWO_GUIDE_MISMATCH;
            goto done;
        }
    }

    /* calcSeqWeights (Tree.cpp:22-100), firstSeq 0, lastSeq n; the reference
     * sets distanceTree=true in readTree and noWeights defaults to false */
    if (T.distance_tree) {
        int *w = (int *)malloc(sizeof(int) * (size_t)(n + 1));
        int sum = 0;
        for (i = 0; i < n; i++) w[i] = calc_weight(&T, i);
        for (i = 0; i < n; i++) sum += w[i];
        if (sum == 0) {
            for (i = 0; i < n; i++) w[i] = 1;
            sum = i;
        }
        for (i = 0; i < n; i++) {
            weights[i] = (w[i] * INT_SCALE_FACTOR) / sum;
            if (weights[i] < 1) weights[i] = 1;
        }
        free(w);
    } else {
        int temp = INT_SCALE_FACTOR / n;
        for (i = 0; i < n; i++) weights[i] = temp;
    }

    /* calcSimilarities (Tree.cpp:327-469): tree path lengths in float,
     * clamped to [0.01, 1.0], then 100 - 100 d into the 1-based matrix */
    {
        tnode **pathToRoot = (tnode **)malloc(sizeof(tnode *) * (size_t)n * 2 + 16);
        float *distToNode = (float *)malloc(sizeof(float) * (size_t)n * 2 + 16);
        double *dmat = (double *)calloc((size_t)n * n, sizeof(double));
        for (i = 0; i < n; i++) {
            int depth = 0;
            float dist = 0.0f;
            tnode *p = T.olptr[i];
            while (p != NULL) {
                pathToRoot[depth] = p;
                dist += p->dist;
                distToNode[depth] = dist;
                p = p->parent;
                depth++;
            }
            for (j = 0; j < i; j++) {
                int found = 0, nn = 0;
                p = T.olptr[j];
                dist = 0.0f;
                while (!found && p->parent != NULL) {
                    for (k = 0; k < depth; k++)
                        if (p->parent == pathToRoot[k]) {
                            found = 1;
                            nn = k;
                        }
                    dist += p->dist;
                    p = p->parent;
                }
                dmat[(size_t)i * n + j] = dist + distToNode[nn - 1];
            }
        }
        for (i = 0; i < n; i++) {
            dmat[(size_t)i * n + i] = 0.0;
            for (j = 0; j < i; j++) {
                double *v = &dmat[(size_t)i * n + j];
                if (*v < 0.01) *v = 0.01;
                if (*v > 1.0) *v = 1.0;
            }
        }
        for (i = 0; i < n; i++) {
            sim[(size_t)(i + 1) * (n + 1) + (i + 1)] = 0.0;
            for (j = 0; j < i; j++) {
                double value = 100.0 - (dmat[(size_t)i * n + j]) * 100.0;
                sim[(size_t)(i + 1) * (n + 1) + (j + 1)] = value;
                sim[(size_t)(j + 1) * (n + 1) + (i + 1)] = value;
            }
        }
        free(dmat);
        free(distToNode);
        free(pathToRoot);
    }

    /* createSets (Tree.cpp:273-322) over 0..n */
    {
        steps_t s;
        int *groups = (int *)calloc((size_t)n + 1, sizeof(int));
        memset(&s, 0, sizeof s);
        s.n = n;
        group_seqs(&T, T.root, groups, n, &s);
        free(groups);
        *steps_out = s.steps;
        *nsteps_out = s.nsteps;
    }
done:
    for (i = 0; i < T.nall; i++) free(T.all[i]);
    free(T.all);
    free(T.nptr);
    free(T.ptrs);
    free(T.lptr);
    free(T.olptr);
    return status;
}
