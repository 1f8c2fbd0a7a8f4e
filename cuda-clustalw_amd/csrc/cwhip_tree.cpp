// cwhip_tree.cpp -- the host stages either side of the progressive alignment:
// the neighbour-joining guide tree and its PHYLIP text, the guide tree read
// back into weights / similarities / alignment steps, and the CLUSTAL text
// writer.  Paths are relative to CPU/clustalw-2.0.11/src of
// coldfunction/CUDA-clustalW.
//
// All tree arithmetic is IEEE double (NJ) or single (TreeNode distances,
// weights) exactly where the reference has it, in the reference's summation
// order; the file must be built without fast-math and with FP contraction off.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <atomic>
#include <functional>
#include <memory>
#include <thread>
#include <unordered_map>
#include <string>
#include <vector>

#include "../../include/cwhip.h"

namespace {

// ---------------------------------------------------------------------------
// NJTree::generateTree (tree/NJTree.cpp:161-751) on a (n+1)^2 symmetric
// matrix, with ClusterTreeOutput's tree description (treeDesc) and branch
// lengths (leftBranch/rightBranch).
// Host threads for the NJ minimum search: T-1 workers spin on a generation
// counter; run(f) calls f(0..T-1) (f(0) on the caller) and returns when all
// are done.
struct SpinPool {
    int nt = 1;
    std::vector<std::thread> th;
    std::atomic<int> gen{0}, done{0};
    std::atomic<bool> stop{false};
    std::function<void(int)> job;
    explicit SpinPool(int n) : nt(std::max(1, n)) {
        for (int t = 1; t < nt; ++t)
            th.emplace_back([this, t]() {
                int seen = 0;
                for (;;) {
                    int g;
                    while ((g = gen.load(std::memory_order_acquire)) == seen) {
                        if (stop.load(std::memory_order_acquire)) return;
                        std::this_thread::yield();
                    }
                    seen = g;
                    job(t);
                    done.fetch_add(1, std::memory_order_acq_rel);
                }
            });
    }
    void run(const std::function<void(int)>& f) {
        job = f;
        done.store(0, std::memory_order_release);
        gen.fetch_add(1, std::memory_order_acq_rel);
        f(0);
        while (done.load(std::memory_order_acquire) < nt - 1) std::this_thread::yield();
    }
    ~SpinPool() {
        stop.store(true, std::memory_order_release);
        for (auto& t : th) t.join();
    }
};

struct NJ {
    int n;
    std::vector<double> d;        // (n+1)^2, SymMatrix semantics
    std::vector<int> desc;        // treeDesc, (n+1)^2
    std::vector<double> left, right;
    bool sym = true;
    double& D(int i, int j) { return d[(size_t)i * (n + 1) + j]; }
    int& DESC(int i, int j) { return desc[(size_t)i * (n + 1) + j]; }
    void set(int i, int j, double v) { D(i, j) = v; D(j, i) = v; }

    void run()
    {
        const int last = n;
        std::vector<double> sumCols(last + 1, 0.0), sumRows(last + 1, 0.0), join(last + 1, 0.0),
            av(last + 1, 0.0);
        // the live nodes form an ascending linked list (node 0 heads it); sums
        // over it in list order fix the doubles
        const int NIL = -1;
        std::vector<int> nxt(last + 2), prv(last + 2);
        for (int i = 0; i <= last; ++i) {
            nxt[i] = i < last ? i + 1 : NIL;
            prv[i] = i > 0 ? i - 1 : NIL;
        }
        for (int i = 1; i <= last; ++i) D(i, i) = av[i] = 0.0;
        for (int j = nxt[0]; j != NIL; j = nxt[j]) {           // :244-274
            double tmp = 0.0;
            int p = nxt[0];
            for (; p != NIL && p < j; p = nxt[p]) tmp = tmp + D(p, j);
            sumRows[j] = tmp;
            tmp = 0.0;
            if (p != NIL && p == j) p = nxt[p];
            for (; p != NIL; p = nxt[p]) tmp += D(j, p);
            sumCols[j] = tmp;
        }
        double fnseqs = (double)last;
        // D(i, j) == D(j, i) for every pair (a DistMatrix is symmetric): the search
        // may then read rows, which are contiguous
        sym = true;
        for (int i = 1; i <= last && sym; ++i)
            for (int j = i + 1; j <= last; ++j)
                if (D(i, j) != D(j, i)) { sym = false; break; }
        std::vector<int> lv(last + 1);
        std::vector<double> dq(last + 1);
        std::unique_ptr<SpinPool> pool;
        if (last > 256)
            pool.reset(new SpinPool((int)std::max<unsigned>(1, std::min<unsigned>(16, std::thread::hardware_concurrency()))));
        desc.assign((size_t)(n + 1) * (n + 1), 0);
        left.assign(n + 2, 0.0);
        right.assign(n + 2, 0.0);
        for (int nc = 1; nc <= last - 3; ++nc) {
            double sumd = 0.0;
            for (int j = nxt[0]; j != NIL; j = nxt[j]) sumd += sumCols[j];
            const double fnseqs2 = fnseqs - 2.0;
            double tmin = 99999.0 * 2.0 * fnseqs2;
            int mini = 0, minj = 0;
            // the pair (ii < jj) of live nodes with the smallest total, first in
            // the reference's scan order (jj ascending, then ii ascending); the
            // scan is split into contiguous jj ranges on host threads and the
            // ranges' winners are taken in order with the same strict test
            int m = 0;
            for (int j = nxt[0]; j != NIL; j = nxt[j]) {
                lv[m] = j;
                dq[m] = sumCols[j] + sumRows[j];
                ++m;
            }
            const int T = (pool && m > 192) ? pool->nt : 1;
            std::vector<double> ctmin(T, tmin);
            std::vector<int> cmi(T, 0), cmj(T, 0), cut(T + 1, m);
            cut[0] = 1;
            if (T > 1) {
                const double W = 0.5 * (double)m * (double)(m - 1);
                double acc = 0.0;
                int t = 1;
                for (int q = 1; q < m && t < T; ++q) {
                    acc += q;
                    if (acc >= W * t / T) cut[t++] = q + 1;
                }
                for (; t < T; ++t) cut[t] = m;
            }
            auto scan = [&](int t) {
                double tm = ctmin[t];
                int bi = 0, bj = 0;
                for (int q = cut[t]; q < cut[t + 1]; ++q) {
                    const int jj = lv[q];
                    const double djq = dq[q];
                    const double* col = sym ? &d[(size_t)jj * (n + 1)] : nullptr;
                    for (int r = 0; r < q; ++r) {
                        const int ii = lv[r];
                        const double dij = sym ? col[ii] : D(ii, jj);
                        const double total = fnseqs2 * dij + 2.0 * sumd - dq[r] - djq;
                        if (total < tm) {
                            tm = total;
                            bi = ii;
                            bj = jj;
                        }
                    }
                }
                ctmin[t] = tm;
                cmi[t] = bi;
                cmj[t] = bj;
            };
            if (T > 1) pool->run(scan);
            else scan(0);
            for (int t = 0; t < T; ++t)
                if (ctmin[t] < tmin) {
                    tmin = ctmin[t];
                    mini = cmi[t];
                    minj = cmj[t];
                }
            double dio = sumCols[mini] + sumRows[mini];
            double djo = sumCols[minj] + sumRows[minj];
            double dmin = D(mini, minj);
            dio = (dio - dmin) / fnseqs2;
            djo = (djo - dmin) / fnseqs2;
            double bi = (dmin + dio - djo) * 0.5;
            double bj = dmin - bi;
            bi = bi - av[mini];
            bj = bj - av[minj];
            const int typei = av[mini] > 0.0 ? 0 : 1;
            const int typej = av[minj] > 0.0 ? 0 : 1;
            if (std::fabs(bi) < 0.0001) bi = 0.0;
            if (std::fabs(bj) < 0.0001) bj = 0.0;
            left[nc] = bi;
            right[nc] = bj;
            for (int i = 1; i <= last; ++i) DESC(nc, i) = 0;
            auto inherit = [&](int node, int type) {
                if (type == 0) {
                    for (int i = nc - 1; i >= 1; --i)
                        if (DESC(i, node) == 1) {
                            for (int j = 1; j <= last; ++j)
                                if (DESC(i, j) == 1) DESC(nc, j) = 1;
                            break;
                        }
                } else {
                    DESC(nc, node) = 1;
                }
            };
            inherit(mini, typei);
            inherit(minj, typej);
            if (dmin <= 0.0) dmin = 0.000001;
            av[mini] = dmin * 0.5;
            fnseqs = fnseqs - 1.0;
            nxt[prv[minj]] = nxt[minj];
            if (nxt[minj] != NIL) prv[nxt[minj]] = prv[minj];
            prv[minj] = nxt[minj] = NIL;
            for (int j = nxt[0]; j != NIL; j = nxt[j]) {
                double tdi = 0.0, tdj = 0.0;
                if (j < minj) {
                    tdj = D(j, minj);
                    sumCols[j] -= tdj;
                } else if (j > minj) {
                    tdj = D(minj, j);
                    sumRows[j] -= tdj;
                }
                if (j < mini) {
                    tdi = D(j, mini);
                    sumCols[j] -= tdi;
                } else if (j > mini) {
                    tdi = D(mini, j);
                    sumRows[j] -= tdi;
                }
                join[j] = (tdj + tdi) * 0.5;
            }
            for (int j = nxt[0]; j != NIL; j = nxt[j]) {
                if (j < mini) {
                    set(j, mini, join[j]);
                    sumCols[j] += join[j];
                } else if (j > mini) {
                    set(mini, j, join[j]);
                    sumRows[j] += join[j];
                }
            }
            double da = 0.0;
            int j = nxt[0];
            for (; j < mini; j = nxt[j]) da = da + join[j];
            sumRows[mini] = da;
            if (j != NIL && j == mini) j = nxt[j];
            da = 0.0;
            for (; j != NIL; j = nxt[j]) da = da + join[j];
            sumCols[mini] = da;
            sumCols[minj] = sumRows[minj] = 0.0;
            for (j = 1; j <= last; ++j) {
                set(minj, j, 0.0);
                join[j] = 0.0;
            }
        }
        // the last three nodes: a trichotomy (:615-741)
        int l[4], nude = 1;
        for (int i = nxt[0]; i != NIL; i = nxt[i]) l[nude++] = i;
        const double b1 = (D(l[1], l[2]) + D(l[1], l[3]) - D(l[2], l[3])) * 0.5;
        const double b2 = D(l[1], l[2]) - b1;
        const double b3 = D(l[1], l[3]) - b1;
        double branch[4] = {0, b1 - av[l[1]], b2 - av[l[2]], b3 - av[l[3]]};
        for (int k = 1; k <= 3; ++k)
            if (std::fabs(branch[k]) < 0.0001) branch[k] = 0.0;
        left[last - 2] = branch[1];
        left[last - 1] = branch[2];
        left[last] = branch[3];
        for (int i = 1; i <= last; ++i) DESC(last - 2, i) = 0;
        for (int i = 1; i <= 3; ++i) {
            if (av[l[i]] > 0.0) {
                for (int k = last - 3; k >= 1; --k)
                    if (DESC(k, l[i]) == 1) {
                        for (int j = 1; j <= last; ++j)
                            if (DESC(k, j) == 1) DESC(last - 2, j) = i;
                        break;
                    }
            } else {
                DESC(last - 2, l[i]) = i;
            }
        }
    }
};

void put_fixed5(std::string& s, double v)
{
    char t[64];
    std::snprintf(t, sizeof t, "%.5f", v);     // `fixed << setprecision(5)`
    s += t;
}

// ClusterTreeOutput::twoWaySplit (tree/ClusterTreeOutput.cpp:101-220)
int two_way_split(NJ& t, std::string& s, const std::vector<std::string>& names, int startRow, int flag)
{
    const int last = t.n;
    if (startRow != last - 2) s += "(\n";
    for (int pass = 0; pass < 2; ++pass) {
        const int want = pass == 0 ? flag : flag;
        int testCol = 0;
        for (int col = 1; col <= last; ++col)
            if (t.DESC(startRow, col) == want) { testCol = col; break; }
        bool single = true;
        int newRow = 0;
        for (int row = startRow - 1; row >= 1; --row)
            if (t.DESC(row, testCol) == 1) { single = false; newRow = row; break; }
        const double br = pass == 0 ? t.left[startRow] : t.right[startRow];
        if (single) {
            t.DESC(startRow, testCol) = 0;
            s += names[testCol];
            if (pass == 0 && startRow == last - 2) return 0;
        } else {
            for (int col = 1; col <= last; ++col)
                if (t.DESC(startRow, col) == 1 && t.DESC(newRow, col) == 1) t.DESC(startRow, col) = 0;
            two_way_split(t, s, names, newRow, 1);
            if (pass == 0 && startRow == last - 2) return newRow;
        }
        s += ":";
        put_fixed5(s, br);
        s += pass == 0 ? ",\n" : ")\n";
    }
    return startRow;
}

// ClusterTreeOutput::printPhylipTree (:45-99): the three subtrees of the
// final trichotomy
std::string phylip(NJ& t, const std::vector<std::string>& nm)
{
    const int n = t.n;
    std::string s = "(\n";
    two_way_split(t, s, nm, n - 2, 1);
    s += ":";
    put_fixed5(s, t.left[n - 2]);
    s += ",\n";
    two_way_split(t, s, nm, n - 2, 2);
    s += ":";
    put_fixed5(s, t.left[n - 1]);
    s += ",\n";
    two_way_split(t, s, nm, n - 2, 3);
    s += ":";
    put_fixed5(s, t.left[n]);
    s += ");\n";
    return s;
}

// ---------------------------------------------------------------------------
// Tree::readTree and what TreeInterface derives from it (tree/Tree.cpp).
constexpr int MAXNAMES = 150;       // general/clustalw.h:164
constexpr int INT_SCALE = 1000;     // INT_SCALE_FACTOR (OS_UNIX)

struct TNode {
    int parent = -1, left = -1, right = -1;
    int leaf = 0, order = 0;
    float dist = 0.0f;
    std::string name;
};

struct TreeReader {
    const char* txt;
    size_t pos = 0, len = 0;
    int ch = 0;
    std::vector<TNode> nodes;
    std::vector<int> ptrs, lptr;
    bool rooted = true, distance_tree = true;

    int get() { if (pos >= len) { ++pos; return -1; } return (unsigned char)txt[pos++]; }
    void skip_space() { int c; do { c = get(); } while (c >= 0 && std::isspace(c)); --pos; }
    int avail() { nodes.emplace_back(); return (int)nodes.size() - 1; }
    void set_info(int p, int parent, int leaf, const std::string& name, float dist)
    {
        TNode& q = nodes[p];
        q.parent = parent;
        q.leaf = leaf;
        q.dist = dist;
        q.order = 0;
        q.name = name;
        if (leaf == 1) { q.left = q.right = -1; }
    }
    void create_node(int p, int parent)
    {
        nodes[p].parent = parent;
        const int a = avail();
        const int b = avail();
        nodes[p].left = a;
        nodes[p].right = b;
    }
    int insert_node(int p)                               // Tree::insertNode
    {
        const int nn = avail();
        create_node(nn, nodes[p].parent);
        nodes[nn].left = p;
        nodes[p].parent = nn;
        set_info(nn, nodes[p].parent, 0, "", 0.0f);
        return nn;
    }
    void create_tree(int ptree, int parent)              // Tree::createTree
    {
        int type;
        std::string name;
        skip_space();
        ch = get();
        if (ch == '(') {
            type = 0;
            ptrs.push_back(ptree);
            create_node(ptree, parent);
            create_tree(nodes[ptree].left, ptree);
            if (ch == ',') {
                create_tree(nodes[ptree].right, ptree);
                if (ch == ',') {
                    ptree = insert_node(ptree);
                    ptrs.push_back(ptree);
                    create_tree(nodes[ptree].right, ptree);
                    rooted = false;
                }
            }
            skip_space();
            ch = get();
        } else {
            type = 1;
            ptrs.push_back(ptree);
            lptr.push_back(ptree);
            name += (char)ch;
            ch = get();
            int i = 1;
            while (ch != ':' && ch != ',' && ch != ')' && ch >= 0) {
                if (i < MAXNAMES) {
                    name += (char)ch;
                    ++i;
                }
                ch = get();
            }
            if (ch != ':') distance_tree = false;
        }
        float dist = 0.0f;
        if (ch == ':') {
            skip_space();
            char* end = nullptr;
            dist = std::strtof(txt + pos, &end);
            pos += (size_t)(end - (txt + pos));
            skip_space();
            ch = get();
        }
        set_info(ptree, parent, type, name, dist);
    }

    // Tree::calcMean.  `pos` (one entry per node, all -1 on entry and on exit)
    // gives a node's position on nptr's root path: a path holds each node once,
    // so it finds the position the reference's linear search finds.
    float calc_mean(int nptr, float* maxDist, int nSeqs, std::vector<int>& pos) const
    {
        float dist = 0.0f, leftSum = 0.0f, rightSum = 0.0f;
        std::vector<int> path;
        std::vector<float> distToNode;
        path.reserve(nSeqs);
        *maxDist = 0.0f;
        for (int p = nptr; p >= 0; p = nodes[p].parent) {
            pos[p] = (int)path.size();
            path.push_back(p);
            dist += nodes[p].dist;
            distToNode.push_back(dist);
        }
        int numLeft = 0, numRight = 0;
        for (int leafp : lptr) {
            int p = leafp, n = 0;
            bool right;
            if (p == nptr) {
                right = true;
                dist = 0.0f;
            } else {
                right = false;
                dist = 0.0f;
                bool found = false;
                while (!found && nodes[p].parent >= 0) {
                    const int k = pos[nodes[p].parent];
                    if (k >= 0) { found = true; n = k; }
                    dist += nodes[p].dist;
                    p = nodes[p].parent;
                }
                if (p == nptr) right = true;
            }
            if (!right) {
                leftSum += dist;
                leftSum += distToNode[n - 1];
                ++numLeft;
            } else {
                rightSum += dist;
                ++numRight;
            }
            if (dist > *maxDist) *maxDist = dist;
        }
        for (int p : path) pos[p] = -1;
        const float leftMean = leftSum / numLeft;
        const float rightMean = rightSum / numRight;
        return leftMean - rightMean;
    }

    int insert_root(int p, float diff)                    // Tree::insertRoot
    {
        const int newp = avail();
        const int t = nodes[p].parent;
        float prevDist = nodes[t].dist;
        nodes[p].parent = newp;
        float dist = nodes[p].dist;
        nodes[p].dist = diff / 2;
        if (nodes[p].dist < 0.0) nodes[p].dist = 0.0f;
        if (nodes[p].dist > dist) nodes[p].dist = dist;
        nodes[t].dist = dist - nodes[p].dist;
        nodes[newp].left = t;
        nodes[newp].right = p;
        nodes[newp].parent = -1;
        nodes[newp].dist = 0.0f;
        nodes[newp].leaf = 0;
        if (nodes[t].left == p) nodes[t].left = nodes[t].parent;
        else nodes[t].right = nodes[t].parent;
        int prev = t, q = nodes[t].parent;
        nodes[t].parent = newp;
        while (q >= 0) {
            if (nodes[q].left == prev) {
                nodes[q].left = nodes[q].parent;
                nodes[q].parent = prev;
                const float td = nodes[q].dist;
                nodes[q].dist = prevDist;
                prevDist = td;
                prev = q;
                q = nodes[q].left;
            } else {
                nodes[q].right = nodes[q].parent;
                nodes[q].parent = prev;
                const float td = nodes[q].dist;
                nodes[q].dist = prevDist;
                prevDist = td;
                prev = q;
                q = nodes[q].right;
            }
        }
        // remove the old root node
        q = prev;
        const int pp = nodes[prev].parent;
        if (nodes[q].left < 0) {
            dist = nodes[q].dist;
            q = nodes[q].right;
            nodes[q].dist += dist;
            nodes[q].parent = pp;
            if (nodes[pp].left == prev) nodes[pp].left = q;
            else nodes[pp].right = q;
            nodes[prev].right = -1;
        } else {
            dist = nodes[q].dist;
            q = nodes[q].left;
            nodes[q].dist += dist;
            nodes[q].parent = pp;
            if (nodes[pp].left == prev) nodes[pp].left = q;
            else nodes[pp].right = q;
            nodes[prev].left = -1;
        }
        return newp;
    }

    int re_root(int ptree, int nseqs)                      // Tree::reRoot
    {
        int rootPtr = ptree;
        float minDiff = 0.0f, minDepth = 1.0f, maxDist = 0.0f;
        bool first = true;
        // every node's (diff, maxDist) on host threads, then the reference's
        // selection in node order
        const size_t np = ptrs.size();
        std::vector<float> diffs(np, 0.0f), maxd(np, 0.0f);
        std::atomic<size_t> nexti{0};
        auto work = [&]() {
            std::vector<int> pos(nodes.size(), -1);
            for (size_t i; (i = nexti.fetch_add(1)) < np;)
                if (nodes[ptrs[i]].parent >= 0) diffs[i] = calc_mean(ptrs[i], &maxd[i], nseqs, pos);
        };
        const int nt = np > 512 ? (int)std::max<unsigned>(1, std::min<unsigned>(16, std::thread::hardware_concurrency())) : 1;
        std::vector<std::thread> th;
        for (int t = 1; t < nt; ++t) th.emplace_back(work);
        work();
        for (auto& t : th) t.join();
        for (size_t i = 0; i < np; ++i) {
            const int p = ptrs[i];
            if (nodes[p].parent < 0) continue;
            const float diff = diffs[i];
            maxDist = maxd[i];
            if (diff == 0 || (diff > 0 && diff < 2 * nodes[p].dist)) {
                if (maxDist < minDepth || first) {
                    first = false;
                    rootPtr = p;
                    minDepth = maxDist;
                    minDiff = diff;
                }
            }
        }
        if (rootPtr == ptree) {
            minDiff = nodes[nodes[rootPtr].left].dist + nodes[nodes[rootPtr].right].dist;
            rootPtr = nodes[rootPtr].right;
        }
        return insert_root(rootPtr, minDiff);
    }
};

std::string fold_name(const std::string& in, bool underscore)
{
    std::string out;
    for (size_t k = 0; k < in.size() && k < (size_t)MAXNAMES; ++k) {
        char c = in[k];
        if (c > 0x40 && c < 0x5b) c = (char)(c | 0x20);
        if (underscore && c == ' ') c = '_';
        out += c;
    }
    return out;
}

} // namespace

extern "C" {

int cwhip_nj_tree(int nseq, const double* dist, const char* const* names, char* out,
                  int64_t cap, int64_t* out_len)
{
    if (nseq < 2 || !dist || !names || !out_len) return CWHIP_E_ARG;
    std::vector<std::string> nm(nseq + 1);
    for (int i = 1; i <= nseq; ++i) {
        if (!names[i - 1]) return CWHIP_E_ARG;
        nm[i] = names[i - 1];
    }
    std::string s;
    if (nseq == 2) {
        // UnRootedClusterTree::treeFromDistMatrix for two sequences
        // (tree/UnRootedClusterTree.cpp:225-233): the float d/2 written with
        // setprecision(5) in the stream's general format
        char t[400];
        const float half = (float)(dist[1 * 3 + 2] / 2.0);
        std::snprintf(t, sizeof t, "(%s:%.5g,%s:%.5g);\n", nm[1].c_str(), (double)half,
                      nm[2].c_str(), (double)half);
        s = t;
    } else {
        NJ t;
        t.n = nseq;
        t.d.assign(dist, dist + (size_t)(nseq + 1) * (nseq + 1));
        t.run();
        s = phylip(t, nm);
    }
    *out_len = (int64_t)s.size();
    if (!out || cap < (int64_t)s.size() + 1) return CWHIP_E_ARG;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return CWHIP_OK;
}

int cwhip_guide_from_tree(int nseq, const char* dnd, const char* const* names, int32_t* weights,
                          double* sim, int32_t* steps, int32_t* nsteps, int32_t* distance_tree)
{
    if (nseq < 2 || !dnd || !names || !weights || !sim || !steps || !nsteps) return CWHIP_E_ARG;
    TreeReader T;
    T.txt = dnd;
    T.len = std::strlen(dnd);
    T.skip_space();
    if (T.get() != '(') return CWHIP_E_ARG;
    T.pos = 0;
    const int seqTree = T.avail();
    T.set_info(seqTree, -1, 0, "", 0.0f);
    T.create_tree(seqTree, -1);
    if ((int)T.lptr.size() != nseq) return CWHIP_E_ARG;
    if (!T.distance_tree && !T.rooted) return CWHIP_E_ARG;
    if (distance_tree) *distance_tree = T.distance_tree ? 1 : 0;
    const int root = T.rooted ? seqTree : T.re_root(seqTree, nseq + 1);
    for (int leafp : T.lptr)                                 // orderNodes
        for (int p = leafp; p >= 0; p = T.nodes[p].parent) T.nodes[p].order++;
    // a sequence's leaf is the last leaf (in tree order) whose folded name is
    // its folded name
    std::vector<int> olptr(nseq, -1);
    std::unordered_map<std::string, int> leaf_of;
    for (int leafp : T.lptr) leaf_of[fold_name(T.nodes[leafp].name, false)] = leafp;
    for (int i = 0; i < nseq; ++i) {
        if (!names[i]) return CWHIP_E_ARG;
        const auto it = leaf_of.find(fold_name(names[i], true));
        if (it == leaf_of.end()) return CWHIP_E_ARG;
        olptr[i] = it->second;
    }
    // calcSeqWeights (Tree.cpp:22-100)
    if (T.distance_tree) {
        std::vector<int> w(nseq);
        int sum = 0;
        for (int i = 0; i < nseq; ++i) {
            float weight = 0.0f;
            for (int p = olptr[i]; T.nodes[p].parent >= 0; p = T.nodes[p].parent)
                weight += T.nodes[p].dist / T.nodes[p].order;
            weight *= 100.0;
            w[i] = (int)weight;
            sum += w[i];
        }
        if (sum == 0) {
            for (int i = 0; i < nseq; ++i) w[i] = 1;
            sum = nseq;
        }
        for (int i = 0; i < nseq; ++i) {
            weights[i] = (w[i] * INT_SCALE) / sum;
            if (weights[i] < 1) weights[i] = 1;
        }
    } else {
        for (int i = 0; i < nseq; ++i) weights[i] = INT_SCALE / nseq;
    }
    // calcSimilarities (Tree.cpp:327-469).  Row i walks each j < i up to the
    // first ancestor on i's root path; a node's position on that path comes
    // from a table instead of the reference's linear search (a path holds each
    // node once, so both find the same position), the float sums keep their
    // order.  Rows run on host threads; each writes only its own pairs.
    {
        const size_t W = (size_t)nseq + 1;
        const int nn = (int)T.nodes.size();
        std::atomic<int> nexti{0};
        auto rows = [&]() {
            std::vector<int> path, pos(nn, -1);
            std::vector<float> distToNode;
            for (int i; (i = nexti.fetch_add(1)) < nseq;) {
                path.clear();
                distToNode.clear();
                float dist = 0.0f;
                for (int p = olptr[i]; p >= 0; p = T.nodes[p].parent) {
                    pos[p] = (int)path.size();
                    path.push_back(p);
                    dist += T.nodes[p].dist;
                    distToNode.push_back(dist);
                }
                sim[(i + 1) * W + (i + 1)] = 0.0;
                for (int j = 0; j < i; ++j) {
                    int p = olptr[j], n = 0;
                    bool found = false;
                    dist = 0.0f;
                    while (!found && T.nodes[p].parent >= 0) {
                        const int k = pos[T.nodes[p].parent];
                        if (k >= 0) { found = true; n = k; }
                        dist += T.nodes[p].dist;
                        p = T.nodes[p].parent;
                    }
                    double dm = dist + distToNode[n - 1];
                    if (dm < 0.01) dm = 0.01;
                    if (dm > 1.0) dm = 1.0;
                    const double value = 100.0 - dm * 100.0;
                    sim[(i + 1) * W + (j + 1)] = value;
                    sim[(j + 1) * W + (i + 1)] = value;
                }
                for (int p : path) pos[p] = -1;
            }
        };
        const int nt = nseq > 256 ? (int)std::max<unsigned>(1, std::min<unsigned>(16, std::thread::hardware_concurrency())) : 1;
        std::vector<std::thread> th;
        for (int t = 1; t < nt; ++t) th.emplace_back(rows);
        rows();
        for (auto& t : th) t.join();
        for (size_t k = 0; k < W; ++k) sim[k] = sim[k * W] = 0.0;
    }
    // createSets / groupSeqs (Tree.cpp:273-322, 1120-1226)
    int ns = 0;
    std::vector<int> next(nseq, 0);
    struct Rec {
        static void go(TreeReader& T, const std::vector<int>& olptr, int p, std::vector<int>& nextGroups,
                       int n, int32_t* steps, int& ns)
        {
            std::vector<int> tmp(n, 0);
            const TNode& q = T.nodes[p];
            if (q.left >= 0) {
                if (T.nodes[q.left].leaf == 0) {
                    go(T, olptr, q.left, nextGroups, n, steps, ns);
                    for (int i = 0; i < n; ++i)
                        if (nextGroups[i] != 0) tmp[i] = 1;
                } else {
                    for (int i = 0; i < n; ++i) tmp[i] = olptr[i] == q.left ? 1 : 0;
                }
            }
            if (q.right >= 0) {
                if (T.nodes[q.right].leaf == 0) {
                    go(T, olptr, q.right, nextGroups, n, steps, ns);
                    for (int i = 0; i < n; ++i)
                        if (nextGroups[i] != 0) tmp[i] = 2;
                } else {
                    for (int i = 0; i < n; ++i) {
                        if (olptr[i] == q.right) tmp[i] = 2;
                        else if (tmp[i] != 0) tmp[i] = 1;
                    }
                }
                int32_t* row = steps + (size_t)ns * (n + 1);
                row[0] = 0;
                for (int i = 0; i < n; ++i) row[i + 1] = tmp[i];
                ++ns;
            }
            nextGroups = tmp;
        }
    };
    Rec::go(T, olptr, root, next, nseq, steps, ns);
    *nsteps = ns;
    return CWHIP_OK;
}

// AlignmentOutput::clustalOut (alignment/AlignmentOutput.cpp:1573-2116) for the
// whole alignment (firstRes 1 .. longest row, all sequences), no sequence
// numbers, no structure masks; conservation groups of the constructor (:39-61).
int cwhip_clustal_text(int nseq, const char* const* names, const uint8_t* rows, int64_t stride,
                       const int32_t* lens, const int32_t* order, int dna, char* out, int64_t cap,
                       int64_t* out_len)
{
    if (nseq < 1 || !names || !rows || !lens || !order || !out_len) return CWHIP_E_ARG;
    static const char* const strong[] = {"STA", "NEQK", "NHQK", "NDEQ", "QHRK",
                                         "MILV", "MILF", "HY", "FYW"};
    static const char* const weak[] = {"CSA", "ATV", "SAG", "STNK", "STPA", "SGND",
                                       "SNDEQK", "NDEQHK", "NEQHRK", "FVLIM", "HFY"};
    static const char kCodes[] = "ABCDEFGHIKLMNOPQRSTUVWXYZ-";
    const int maxAA = 24, nStrong = 9, nWeak = 11;
    // per code 0..25: bit k set when the code's letter is in group k
    uint32_t strongBits[26], weakBits[26];
    for (int c = 0; c < 26; ++c) {
        strongBits[c] = weakBits[c] = 0;
        for (int k = 0; k < nStrong; ++k)
            if (std::strchr(strong[k], kCodes[c])) strongBits[c] |= 1u << k;
        for (int k = 0; k < nWeak; ++k)
            if (std::strchr(weak[k], kCodes[c])) weakBits[c] |= 1u << k;
    }
    int maxNames = 0, length = 0;
    for (int i = 0; i < nseq; ++i) {
        if (!names[i] || lens[i] < 0 || lens[i] > stride) return CWHIP_E_ARG;
        maxNames = std::max(maxNames, (int)std::strlen(names[i]));
        length = std::max(length, (int)lens[i]);
    }
    for (int k = 0; k < nseq; ++k)
        if (order[k] < 1 || order[k] > nseq) return CWHIP_E_ARG;
    if (maxNames < 10) maxNames = 10;                            // Alignment.cpp:794-796
    const int namesWidth = std::max(std::min(30, maxNames), 10);  // MAXNAMESTODISPLAY / MIN...
    int lineLength = 80 - namesWidth;                            // PAGEWIDTH
    lineLength = lineLength - lineLength % 10;
    if (lineLength > 60 || lineLength <= 0) lineLength = 60;     // LINELENGTH
    int chunks = length / lineLength;
    if (length % lineLength != 0) ++chunks;
    auto at = [&](int seq, int col) -> int {                    // 1-based sequence and column
        return col <= lens[seq - 1] ? rows[(int64_t)(seq - 1) * stride + col - 1] : -3;
    };
    // each name cut or padded to namesWidth, then 6 blanks
    std::vector<std::string> label(nseq);
    for (int k = 0; k < nseq; ++k) {
        label[k] = names[k];
        if ((int)label[k].size() > namesWidth) label[k].resize(namesWidth);
        label[k].resize(namesWidth + 6, ' ');
    }
    std::string s = "CLUSTAL 2.0.11 multiple sequence alignment\n\n";
    s.reserve(s.size() + (size_t)chunks * (1 + (size_t)(nseq + 1) * (namesWidth + 7 + lineLength)));
    std::string line;
    for (int lv1 = 1; lv1 <= chunks; ++lv1) {
        const int pos = (lv1 - 1) * lineLength + 1;
        const int ptr = std::min(length, pos + lineLength - 1);
        s += "\n";
        for (int ii = 1; ii <= nseq; ++ii) {
            const int i = order[ii - 1];
            s += label[i - 1];
            const size_t at0 = s.size();
            s.append(ptr - pos + 1, '-');
            char* o = &s[at0];
            const uint8_t* r = rows + (int64_t)(i - 1) * stride;
            const int end = std::min(ptr, (int)lens[i - 1]);        // past a row's end: '-'
            for (int j = pos; j <= end; ++j) {
                const int val = r[j - 1];
                if (val == 253) break;
                o[j - pos] = val > maxAA ? '-' : kCodes[val];
            }
            s += '\n';
        }
        // A column is '*' when every sequence counts and equals the first, ':'
        // when every sequence counts and has a residue of one strong group,
        // '.' likewise for a weak group; a sequence counts when both it and
        // the first reach the column and the first's code is a residue.  So a
        // group's count reaches nseq exactly when the AND over the sequences
        // of their codes' group bits keeps its bit.
        line.assign(ptr - pos + 1, ' ');
        for (int i = pos; i <= ptr; ++i) {
            const int r1 = i <= lens[0] ? at(1, i) : -1;
            if (r1 < 0 || r1 > maxAA) continue;
            bool ident = true;
            uint32_t sAnd = ~0u, wAnd = ~0u;
            for (int j = 1; j <= nseq && (ident || sAnd || wAnd); ++j) {
                if (i > lens[j - 1]) {
                    ident = false;
                    sAnd = wAnd = 0;
                    break;
                }
                const int rj = at(j, i);
                ident = ident && rj == r1;
                sAnd &= rj <= maxAA + 1 ? strongBits[rj] : 0u;
                wAnd &= rj <= maxAA + 1 ? weakBits[rj] : 0u;
            }
            char mark = ' ';
            if (ident) mark = '*';
            else if (!dna && sAnd) mark = ':';
            else if (!dna && wAnd) mark = '.';
            line[i - pos] = mark;
        }
        s += std::string(namesWidth + 6, ' ');
        s += line;
        s += "\n";
    }
    *out_len = (int64_t)s.size();
    if (!out || cap < (int64_t)s.size() + 1) return CWHIP_E_ARG;
    std::memcpy(out, s.c_str(), s.size() + 1);
    return CWHIP_OK;
}

}  // extern "C"

std::string cw_phylip_text(int n, std::vector<int32_t>& desc, const std::vector<double>& left,
                           const std::vector<double>& right, const char* const* names)
{
    NJ t;
    t.n = n;
    t.desc.swap(desc);
    t.left = left;
    t.right = right;
    std::vector<std::string> nm(n + 1);
    for (int i = 1; i <= n; ++i) nm[i] = names[i - 1];
    std::string s = phylip(t, nm);
    t.desc.swap(desc);
    return s;
}
