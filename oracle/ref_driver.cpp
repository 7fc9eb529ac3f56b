// C entry points over the reference's own host functions (read_network,
// get_triples, get_dec_digits, report_name: large_strain_relax.cu:42-213),
// compiled from the reference's source by `make -C oracle ref` into
// oracle/_ref/libref_host.so.  TEST INFRASTRUCTURE ONLY: tests/test_ref.py
// compares the oracle's and the product's readers, triples and report names
// with these, bit for bit.
#include <cstdio>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

using namespace std;

// the reference's element records (large_strain_relax.cu:33-34)
typedef tuple<int, int, double, char> edge_datum;
typedef tuple<int, int, int, char, char> triple_datum;

// defined by the reference's translation unit
bool read_network(string file_name, double &offset, vector<double> &points,
                  vector<edge_datum> &edges, double &min_y, double &max_y);
vector<triple_datum> get_triples(vector<double> points, vector<edge_datum> edges, double offset);
int get_dec_digits(int value);
string report_name(string base, int num_digits, int count, string ext);

namespace {
struct RefNet {
    double offset = 0, miny = 0, maxy = 0;
    vector<double> points;
    vector<edge_datum> edges;
    vector<triple_datum> triples;
};
}  // namespace

extern "C" {

// read_network, then get_triples when triples != 0; NULL when the reader fails.
void *ref_load(const char *path, int triples)
{
    RefNet *r = new RefNet;
    if (!read_network(path, r->offset, r->points, r->edges, r->miny, r->maxy)) {
        delete r;
        return nullptr;
    }
    if (triples) r->triples = get_triples(r->points, r->edges, r->offset);
    return r;
}

void ref_sizes(const void *h, int *n, int *e, int *t, double *info)
{
    const RefNet *r = static_cast<const RefNet *>(h);
    *n = (int)(r->points.size() / 2);
    *e = (int)r->edges.size();
    *t = (int)r->triples.size();
    info[0] = r->offset;
    info[1] = r->miny;
    info[2] = r->maxy;
}

// pos[2n]; bonds[2e], rest[e], wrap[e]; tri[3t], twrap[2t]
void ref_arrays(const void *h, double *pos, int *bonds, double *rest, signed char *wrap, int *tri,
                signed char *twrap)
{
    const RefNet *r = static_cast<const RefNet *>(h);
    std::memcpy(pos, r->points.data(), sizeof(double) * r->points.size());
    for (size_t k = 0; k < r->edges.size(); k++) {
        bonds[2 * k] = get<0>(r->edges[k]);
        bonds[2 * k + 1] = get<1>(r->edges[k]);
        rest[k] = get<2>(r->edges[k]);
        wrap[k] = (signed char)get<3>(r->edges[k]);
    }
    for (size_t k = 0; k < r->triples.size(); k++) {
        tri[3 * k] = get<0>(r->triples[k]);
        tri[3 * k + 1] = get<1>(r->triples[k]);
        tri[3 * k + 2] = get<2>(r->triples[k]);
        twrap[2 * k] = (signed char)get<3>(r->triples[k]);
        twrap[2 * k + 1] = (signed char)get<4>(r->triples[k]);
    }
}

void ref_free(void *h) { delete static_cast<RefNet *>(h); }

int ref_dec_digits(int v) { return get_dec_digits(v); }

// report_name into out (cap bytes); the name's length
int ref_report_name(const char *base, int digits, int count, const char *ext, char *out, int cap)
{
    const string s = report_name(base, digits, count, ext);
    std::snprintf(out, (size_t)cap, "%s", s.c_str());
    return (int)s.size();
}

}  // extern "C"
