// Points, point containers, Voronoi cells, complexes and diagrams
// (include/point.h, point.hpp) for one process: containers own host atoms as
// in the reference and keep a device copy (rnng_dataset_t for float points,
// rnng_strings_t for char strings) created on first use.
#ifndef RNNG_HOST_POINT_H_
#define RNNG_HOST_POINT_H_

#include <fstream>
#include <memory>

#include "simplex.h"

template <class Atom_>
class Point {
public:
    using Atom = Atom_;
    Point(const Atom* mem, Index dim, Index index) : mem(mem), dim(dim), index(index), container_offset(-1) {}
    Point(const Atom* mem, Index dim, Index index, Index container_offset)
        : mem(mem), dim(dim), index(index), container_offset(container_offset) {}
    Index id() const { return index; }
    Index size() const { return dim; }
    Index offset() const { return container_offset; }
    Atom operator[](Index i) const { return mem[i]; }
    explicit operator const Atom*() const { return mem; }
    const Atom* begin() const { return mem; }
    const Atom* end() const { return mem + dim; }

private:
    const Atom* mem;
    Index dim, index, container_offset;
};

namespace rnng_host {

struct DatasetHandle {
    rnng_dataset_t h = nullptr;
    ~DatasetHandle() {
        if (h) rnng_dataset_destroy(h);
    }
};
struct StringsHandle {
    rnng_strings_t h = nullptr;
    ~StringsHandle() {
        if (h) rnng_strings_destroy(h);
    }
};
struct GraphOwner {
    rnng_graph_t h = nullptr;
    ~GraphOwner() {
        if (h) rnng_graph_destroy(h);
    }
};
struct FiltrationOwner {
    rnng_filtration_t h = nullptr;
    ~FiltrationOwner() {
        if (h) rnng_filtration_destroy(h);
    }
};

// (i, j, w) of a device graph, sorted by (i, j).
inline void graph_edges(rnng_graph_t g, std::vector<int64_t>& ij, std::vector<float>& w) {
    int64_t ne = 0;
    check(rnng_graph_num_edges(g, &ne), "graph_num_edges");
    ij.resize((size_t)(2 * ne));
    w.resize((size_t)ne);
    if (ne > 0) check(rnng_graph_copy(g, ij.data(), w.data(), ne, 1), "graph_copy");
}

// The pairs within the radius of a diagram's batched cell build that one
// cell's own queries return (drivers/landmark.cpp:203-210): member pairs
// (local offsets, each pair once) and (member, ghost) pairs, with their float
// distances.  A query of the cell against itself or its ghosts at that radius
// is answered from here instead of being computed again.
struct CellPairs {
    Real radius = -1;
    std::vector<std::tuple<int32_t, int32_t, float>> members, ghosts;
};

}  // namespace rnng_host

template <class Atom_>
class PointContainer {
public:
    using Atom = Atom_;
    using AtomVector = std::vector<Atom>;

    PointContainer() : offsets({0}) {}
    explicit PointContainer(const std::vector<Point<Atom>>& points) : offsets({0}) {
        for (const auto& p : points) push_back(p);
    }
    // points of equal dimension `dim`, ids given
    PointContainer(const AtomVector& atoms, Index size, Index dim, const IndexVector& indices)
        : data(atoms), ids(indices) {
        offsets.resize((size_t)size + 1);
        for (Index i = 0; i <= size; ++i) offsets[(size_t)i] = i * dim;
    }
    // variable-length points: sizes[i] atoms each
    PointContainer(const AtomVector& atoms, const IndexVector& sizes, const IndexVector& indices)
        : data(atoms), ids(indices) {
        offsets.assign(1, 0);
        for (Index s : sizes) offsets.push_back(offsets.back() + s);
    }
    // the points of lhs followed by those of rhs (include/point.h:47: a cell
    // and its ghosts as one complex)
    PointContainer(const PointContainer& lhs, const PointContainer& rhs)
        : data(lhs.data), offsets(lhs.offsets), ids(lhs.ids) {
        data.insert(data.end(), rhs.data.begin(), rhs.data.end());
        const Index base = lhs.num_atoms();
        for (size_t k = 1; k < rhs.offsets.size(); ++k) offsets.push_back(base + rhs.offsets[k]);
        ids.insert(ids.end(), rhs.ids.begin(), rhs.ids.end());
    }

    Index num_points() const { return (Index)ids.size(); }
    Index num_atoms() const { return (Index)data.size(); }
    Point<Atom> operator[](Index i) const { return Point<Atom>(mem(i), size(i), id(i), i); }

    // .fbin / .u8bin body: uint32 n, uint32 d, then n*d float32 or uint8.
    Index read_bin(const char* fname, bool u8) {
        static_assert(std::is_same<Atom, float>::value, "binary point files hold float or u8 points");
        std::ifstream is(fname, std::ios::binary);
        if (!is) fail(std::string("cannot open ") + fname);
        uint32_t h[2] = {0, 0};
        if (!is.read(reinterpret_cast<char*>(h), sizeof(h))) fail(std::string("missing header in ") + fname);
        const size_t n = h[0], d = h[1], m = n * d;
        data.resize(m);
        if (u8) {
            std::vector<unsigned char> b(m);
            if (m && !is.read(reinterpret_cast<char*>(b.data()), (std::streamsize)m)) fail(std::string("short file ") + fname);
            for (size_t i = 0; i < m; ++i) data[i] = (float)b[i];
        } else if (m && !is.read(reinterpret_cast<char*>(data.data()), (std::streamsize)(m * sizeof(float)))) {
            fail(std::string("short file ") + fname);
        }
        offsets.resize(n + 1);
        ids.resize(n);
        for (size_t i = 0; i <= n; ++i) offsets[i] = (Index)(i * d);
        for (size_t i = 0; i < n; ++i) ids[i] = (Index)i;
        invalidate();
        return (Index)n;
    }

    // fvecs records: int32 dimension then that many float32 (include/point.hpp:112-168).
    Index read_fvecs(const char* fname) {
        static_assert(std::is_same<Atom, float>::value, "read_fvecs reads float points");
        std::ifstream is(fname, std::ios::binary);
        if (!is) fail(std::string("cannot open ") + fname);
        is.seekg(0, is.end);
        const size_t filesize = (size_t)is.tellg();
        is.seekg(0, is.beg);
        int d = 0;
        if (filesize) is.read((char*)&d, sizeof(int));
        is.seekg(0, is.beg);
        const size_t rec = sizeof(int) + sizeof(float) * (size_t)d;
        if (filesize && (d <= 0 || filesize % rec != 0)) fail(std::string("malformed fvecs file ") + fname);
        const size_t total = filesize ? filesize / rec : 0;
        data.resize(total * (size_t)d);
        offsets.resize(total + 1);
        ids.resize(total);
        std::vector<char> p(rec);
        for (size_t i = 0; i < total; ++i) {
            is.read(p.data(), (std::streamsize)rec);
            int dt;
            std::memcpy(&dt, p.data(), sizeof(int));
            if (dt != d) fail(std::string("mixed dimensions in ") + fname);
            std::memcpy(&data[i * (size_t)d], p.data() + sizeof(int), sizeof(float) * (size_t)d);
            offsets[i] = (Index)(i * (size_t)d);
            ids[i] = (Index)i;
        }
        offsets[total] = (Index)(total * (size_t)d);
        invalidate();
        return (Index)total;
    }

    // .fbin / .u8bin (uint32 n, uint32 d, then n*d float32 or uint8;
    // dataset_io.py:48-60): u8 values become floats.
    Index read_fbin(const char* fname) { return read_bin(fname, false); }
    Index read_u8bin(const char* fname) { return read_bin(fname, true); }

    // By extension as dataset_io.py's read_file: .fbin, .u8bin; anything else
    // is read as fvecs, as the reference drivers read every input.
    Index read_file(const char* fname) {
        const std::string f(fname);
        auto ends = [&](const char* e) {
            const std::string x(e);
            return f.size() >= x.size() && f.compare(f.size() - x.size(), x.size(), x) == 0;
        };
        if (ends(".fbin")) return read_fbin(fname);
        if (ends(".u8bin")) return read_u8bin(fname);
        return read_fvecs(fname);
    }

    // One string per non-empty line, a trailing '\r' dropped (include/point.hpp:170-205).
    Index read_seqs(const char* fname) {
        static_assert(std::is_same<Atom, char>::value, "read_seqs reads strings");
        std::ifstream is(fname);
        if (!is) fail(std::string("cannot open ") + fname);
        offsets.clear();
        ids.clear();
        data.clear();
        std::string line;
        Index id = 0;
        while (std::getline(is, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            if (line.empty()) continue;
            offsets.push_back((Index)data.size());
            ids.push_back(id++);
            data.insert(data.end(), line.begin(), line.end());
        }
        offsets.push_back((Index)data.size());
        invalidate();
        return id;
    }

    // The communicator forms (include/point.hpp:207-367): one process reads
    // the whole file and returns the total count.
    Index read_fvecs(const char* fname, SingleComm) { return read_fvecs(fname); }
    Index read_seqs(const char* fname, SingleComm) { return read_seqs(fname); }

    // The systolic ring's exchange (include/point.hpp:369-417): with one rank
    // the block travels to the same rank, so the receive buffer gets a copy.
    struct SendrecvRequest {
        void wait() {}
    };
    void sendrecv(PointContainer& recvbuf, int /*recvrank*/, int /*sendrank*/, SingleComm, SendrecvRequest&) const {
        recvbuf = *this;
    }
    // Every rank's points (include/point.hpp:427-486): one rank's are all.
    void allgather(const PointContainer& sendbuf, SingleComm) { *this = sendbuf; }

    void push_back(const Point<Atom>& p) {
        data.insert(data.end(), p.begin(), p.end());
        offsets.push_back((Index)data.size());
        ids.push_back(p.id());
        invalidate();
    }
    void indexed_gather(const PointContainer& src, const IndexVector& index_offsets) {
        data.clear();
        ids.clear();
        offsets.assign(1, 0);
        for (Index i : index_offsets) push_back(src[i]);
    }
    void swap(PointContainer& o) {
        data.swap(o.data);
        offsets.swap(o.offsets);
        ids.swap(o.ids);
        dev_f.swap(o.dev_f);
        dev_s.swap(o.dev_s);
    }
    typename IndexVector::const_iterator ids_begin() const { return ids.begin(); }
    typename IndexVector::const_iterator ids_end() const { return ids.end(); }

    // Common dimension of float points (aborts if the points differ in size).
    Index dim() const {
        const Index n = num_points();
        if (n == 0) return 0;
        const Index d = size(0);
        for (Index i = 1; i < n; ++i)
            if (size(i) != d) fail("points of different dimensions");
        return d;
    }
    const AtomVector& atoms() const { return data; }
    const IndexVector& atom_offsets() const { return offsets; }

    // Device copies (built once, dropped when the container changes).
    rnng_dataset_t device_points() const {
        static_assert(std::is_same<Atom, float>::value, "device points are float");
        if (!dev_f) {
            auto h = std::make_shared<rnng_host::DatasetHandle>();
            rnng_host::check(rnng_dataset_create_f32(data.data(), num_points(), (int32_t)std::max<Index>(dim(), 1), &h->h),
                             "dataset_create_f32");
            dev_f = h;
        }
        return dev_f->h;
    }
    rnng_strings_t device_strings() const {
        static_assert(std::is_same<Atom, char>::value, "device strings are char");
        if (!dev_s) {
            auto h = std::make_shared<rnng_host::StringsHandle>();
            rnng_host::check(rnng_strings_create(data.data(), offsets.data(), num_points(), &h->h),
                             "strings_create");
            dev_s = h;
        }
        return dev_s->h;
    }

protected:
    AtomVector data;
    IndexVector offsets, ids;
    mutable std::shared_ptr<rnng_host::DatasetHandle> dev_f;
    mutable std::shared_ptr<rnng_host::StringsHandle> dev_s;

    void invalidate() {
        dev_f.reset();
        dev_s.reset();
    }
    [[noreturn]] static void fail(const std::string& msg) { rnng_host::fail(msg); }
    const Atom* mem(Index i) const { return data.data() + offsets[(size_t)i]; }
    Index size(Index i) const { return offsets[(size_t)i + 1] - offsets[(size_t)i]; }
    Index id(Index i) const {
        if (!(i >= 0 && i < (Index)ids.size())) fail("point index out of range");
        return ids[(size_t)i];
    }
};

// A Voronoi cell: its members (center first), ghost points and the members'
// interior flags (include/point.h:102-137).
template <class Atom_>
class VoronoiCell : public PointContainer<Atom_> {
public:
    using Atom = Atom_;
    VoronoiCell(const std::vector<Point<Atom>>& points, const RealVector& dist_to_centers, Index cell_index)
        : PointContainer<Atom>(points), cell_index(cell_index), dist_to_centers(dist_to_centers),
          interior(points.size(), false) {}
    void add_ghost_point(const Point<Atom>& p) {
        ghost_points.push_back(p);
        served.reset();
    }
    Index id() const { return cell_index; }
    Index num_ghosts() const { return ghost_points.num_points(); }
    Point<Atom> ghost(Index i) const { return ghost_points[i]; }
    Real dist_to_center(Index i) const { return dist_to_centers[(size_t)i]; }
    const PointContainer<Atom>& ghosts() const { return ghost_points; }
    const BoolVector& interiors() const { return interior; }
    void set_interior(Index offset) { interior[(size_t)offset] = true; }

    // The pairs of this cell within the radius of the diagram's batched build
    // (null when there is none or the cell changed since).
    const rnng_host::CellPairs* served_pairs() const { return served.get(); }
    void serve_pairs(std::shared_ptr<const rnng_host::CellPairs> p) { served = std::move(p); }

protected:
    Index cell_index;
    RealVector dist_to_centers;
    BoolVector interior;
    PointContainer<Atom> ghost_points;
    std::shared_ptr<const rnng_host::CellPairs> served;
};

// A cell with its ghosts as one point set and the Vietoris-Rips filtration of
// that set (include/point.h:140-170, point.hpp:1126-1290): the clique
// enumeration runs on the GPU over the set's epsilon graph; simplices carry
// the global vertex ids and the interior flag (any vertex a member that is no
// other cell's ghost).  The flags of ghost vertices are false: the reference
// reads them past the end of the members' flags (point.hpp:1132, :1166).
template <class Atom_>
class VoronoiComplex : public PointContainer<Atom_> {
public:
    using Atom = Atom_;

    VoronoiComplex(const VoronoiCell<Atom>& cell, Real radius, Index maxdim)
        : PointContainer<Atom>(cell, cell.ghosts()), radius(radius), maxdim(maxdim), local(cell.num_points()),
          interior(cell.interiors()) {
        interior.resize((size_t)this->num_points(), false);
    }

    template <class Distance>
    void build_filtration(Distance&, Index /*num_verts*/, Real /*cover*/, Index /*leaf_size*/) {
        static_assert(rnng_host::is_l2<Distance>() && std::is_same<Atom, float>::value,
                      "the GPU Rips path computes L2Distance<float>");
        filtration.clear();
        const Index n = this->num_points();
        if (n == 0) return;
        rnng_host::FiltrationOwner f;
        rnng_host::check(rnng_rips_build(this->device_points(), radius, (int32_t)maxdim, &f.h, nullptr), "rips_build");
        IndexVector gids(this->ids_begin(), this->ids_end());
        std::vector<uint8_t> vflag((size_t)n);
        for (Index v = 0; v < n; ++v) vflag[(size_t)v] = interior[(size_t)v] ? 1 : 0;
        int64_t ns = 0;
        rnng_host::check(rnng_filtration_size(f.h, &ns, nullptr), "filtration_size");
        std::vector<uint8_t> sflag((size_t)ns);
        rnng_host::check(rnng_filtration_relabel(f.h, gids.data(), n, vflag.data(), sflag.data()), "filtration_relabel");
        std::vector<float> val((size_t)ns);
        std::vector<int64_t> sid((size_t)ns);
        rnng_host::check(rnng_filtration_copy(f.h, val.data(), sid.data(), ns), "filtration_copy");
        filtration.reserve((size_t)ns);
        for (int64_t s = 0; s < ns; ++s) {
            WeightedSimplex w(sid[(size_t)s], val[(size_t)s]);
            w.interior = sflag[(size_t)s] != 0;
            filtration.push_back(w);
        }
    }

    // "%f\t%lld\t%d" (or the vertex list) per simplex (include/point.hpp:1269-1290).
    void write_filtration_file(const char* fname, Index num_verts, bool use_ids = true) const {
        FILE* f = std::fopen(fname, "w");
        if (!f) rnng_host::fail(std::string("cannot open ") + fname);
        for (const auto& s : filtration) {
            if (use_ids) std::fprintf(f, "%f\t%lld\t%d\n", s.value, (long long)s.getid(), (int)s.interior);
            else std::fprintf(f, "%f\t%s\t%d\n", s.value, s.get_simplex_repr(num_verts).c_str(), (int)s.interior);
        }
        std::fclose(f);
    }

    Index num_simplices() const { return (Index)filtration.size(); }
    typename std::vector<WeightedSimplex>::const_iterator begin() const { return filtration.begin(); }
    typename std::vector<WeightedSimplex>::const_iterator end() const { return filtration.end(); }

private:
    Real radius;
    Index maxdim;
    Index local;
    BoolVector interior;
    std::vector<WeightedSimplex> filtration;
};

// The Voronoi diagram of float points around centers taken from the points
// (the drivers gather their centers by id, drivers/landmark.cpp:111-124):
// exact assignment with ties to the later center (include/point.hpp:535-559),
// cells in coalesce order, ghosts and the epsilon graph through the cells on
// the GPU.
template <class Atom_>
class VoronoiDiagram : public PointContainer<Atom_> {
public:
    using Atom = Atom_;
    using Cell = VoronoiCell<Atom>;
    static_assert(std::is_same<Atom, float>::value, "the GPU Voronoi path is for float points");

    template <class Distance>
    VoronoiDiagram(const PointContainer<Atom>& points, const PointContainer<Atom>& centers, Distance&)
        : PointContainer<Atom>(points) {
        static_assert(rnng_host::is_l2<Distance>(), "the GPU Voronoi path computes L2Distance<float>");
        IndexVector cid(centers.ids_begin(), centers.ids_end());
        // center ids are positions in `points` (the whole set, ids 0..n-1)
        handle_.reset(new Handle());
        rnng_host::check(rnng_voronoi_create(this->device_points(), cid.data(), (int64_t)cid.size(), &handle_->h),
                         "voronoi_create");
        num_centers = (Index)cid.size();
        cell_indices.resize((size_t)points.num_points());
        dist_to_centers.resize((size_t)points.num_points());
        rnng_host::check(rnng_voronoi_assignment(handle_->h, cell_indices.data(), dist_to_centers.data()),
                         "voronoi_assignment");
    }

    // Cells in cell order with members in coalesce order (include/point.hpp:561-758).
    void coalesce_cells(std::vector<Cell>& cells) const {
        IndexVector ptr((size_t)num_centers + 1), ids((size_t)this->num_points());
        rnng_host::check(rnng_voronoi_cells(handle_->h, ptr.data(), ids.data()), "voronoi_cells");
        cells.clear();
        for (Index c = 0; c < num_centers; ++c) {
            std::vector<Point<Atom>> pts;
            RealVector dist;
            for (Index k = ptr[(size_t)c]; k < ptr[(size_t)c + 1]; ++k) {
                pts.push_back((*this)[ids[(size_t)k]]);
                dist.push_back(dist_to_centers[(size_t)ids[(size_t)k]]);
            }
            cells.emplace_back(pts, dist, c);
        }
    }
    void coalesce_cells(std::vector<Cell>& cells, SingleComm) const { coalesce_cells(cells); }

    // Ghost points of every cell (include/point.hpp:760-929), found by one
    // batched build of the epsilon graph through the cells, which then also
    // answers each cell's queries against itself and its ghosts.
    template <class Distance>
    void add_ghost_points_systolic(std::vector<Cell>& cells, Distance&, Real radius, Real /*cover*/,
                                   Index /*leaf_size*/) const {
        static_assert(rnng_host::is_l2<Distance>(), "the GPU Voronoi path computes L2Distance<float>");
        attach_ghosts(cells, radius);
    }
    template <class Distance>
    void add_ghost_points_systolic(std::vector<Cell>& cells, Distance& d, Real radius, Real cover, Index leaf,
                                   SingleComm) const {
        add_ghost_points_systolic(cells, d, radius, cover, leaf);
    }

    // The same ghosts, and each cell's members that are no other cell's ghost
    // marked interior (include/point.hpp:931-1124).
    template <class Distance>
    void add_ghost_points_systolic_rips(std::vector<Cell>& cells, Distance&, Real radius, Real /*cover*/,
                                        Index /*leaf_size*/) const {
        static_assert(rnng_host::is_l2<Distance>(), "the GPU Voronoi path computes L2Distance<float>");
        const std::vector<int64_t> pairs = attach_ghosts(cells, radius);
        std::vector<int32_t> shared((size_t)this->num_points(), 0);
        for (size_t g = 0; g < pairs.size() / 2; ++g) shared[(size_t)pairs[2 * g + 1]]++;
        for (Cell& cell : cells)
            for (Index j = 0; j < cell.num_points(); ++j)
                if (shared[(size_t)cell[j].id()] == 0) cell.set_interior(j);
    }
    template <class Distance>
    void add_ghost_points_systolic_rips(std::vector<Cell>& cells, Distance& d, Real radius, Real cover, Index leaf,
                                        SingleComm) const {
        add_ghost_points_systolic_rips(cells, d, radius, cover, leaf);
    }

    // Undirected edges (i < j, weight = reference float distance) of the graph
    // the landmark driver builds through the cells (drivers/landmark.cpp:185-210).
    std::vector<std::tuple<Index, Index, Real>> epsilon_graph(Real radius, rnng_stats* stats = nullptr) const {
        build(radius, stats);
        std::vector<int64_t> ij;
        std::vector<float> w;
        rnng_host::graph_edges(graph_->h, ij, w);
        std::vector<std::tuple<Index, Index, Real>> out(w.size());
        for (size_t e = 0; e < w.size(); ++e) out[e] = std::make_tuple(ij[2 * e], ij[2 * e + 1], w[e]);
        return out;
    }

    rnng_voronoi_t handle() const { return handle_->h; }

    // The epsilon graph through the cells (device edge list, owned by the
    // diagram), built once per radius.
    rnng_graph_t graph_handle(Real radius) const {
        build(radius);
        return graph_->h;
    }

private:
    struct Handle {
        rnng_voronoi_t h = nullptr;
        ~Handle() {
            if (h) rnng_voronoi_destroy(h);
        }
    };
    void build(Real radius, rnng_stats* stats = nullptr) const {
        if (graph_ && built_radius == radius && !stats) return;
        graph_.reset(new rnng_host::GraphOwner());
        rnng_host::check(rnng_voronoi_graph_build(handle_->h, radius, RNNG_MODE_MFMA, &graph_->h, stats),
                         "voronoi_graph_build");
        built_radius = radius;
    }

    // Ghosts of the batched build added to the cells (cells in coalesce
    // order, as coalesce_cells gives them) and each cell's pairs attached for
    // its queries.  Returns the (cell, point) ghost pairs.
    std::vector<int64_t> attach_ghosts(std::vector<Cell>& cells, Real radius) const {
        if ((Index)cells.size() != num_centers) rnng_host::fail("add_ghost_points: cells are not this diagram's");
        build(radius);
        int64_t ng = 0;
        rnng_host::check(rnng_voronoi_num_ghosts(handle_->h, &ng), "voronoi_num_ghosts");
        std::vector<int64_t> pairs((size_t)(2 * ng));
        if (ng > 0) rnng_host::check(rnng_voronoi_ghosts(handle_->h, pairs.data(), ng), "voronoi_ghosts");
        // ghost lists per cell (sorted by point) and member offsets per point
        std::vector<size_t> gptr((size_t)num_centers + 1, 0);
        for (int64_t g = 0; g < ng; ++g) gptr[(size_t)pairs[(size_t)(2 * g)] + 1]++;
        for (Index c = 0; c < num_centers; ++c) gptr[(size_t)c + 1] += gptr[(size_t)c];
        for (int64_t g = 0; g < ng; ++g)
            cells[(size_t)pairs[(size_t)(2 * g)]].add_ghost_point((*this)[pairs[(size_t)(2 * g + 1)]]);
        std::vector<int32_t> pos((size_t)this->num_points(), -1);
        for (const Cell& cell : cells)
            for (Index j = 0; j < cell.num_points(); ++j) pos[(size_t)cell[j].id()] = (int32_t)j;
        auto ghost_index = [&](Index c, Index p) -> int32_t {
            const int64_t* lo = pairs.data();
            size_t a = gptr[(size_t)c], b = gptr[(size_t)c + 1];
            while (a < b) {
                const size_t mid = (a + b) / 2;
                if (lo[2 * mid + 1] < p) a = mid + 1;
                else b = mid;
            }
            return (a < gptr[(size_t)c + 1] && lo[2 * a + 1] == p) ? (int32_t)(a - gptr[(size_t)c]) : -1;
        };
        std::vector<std::shared_ptr<rnng_host::CellPairs>> cp((size_t)num_centers);
        for (auto& x : cp) {
            x = std::make_shared<rnng_host::CellPairs>();
            x->radius = radius;
        }
        std::vector<int64_t> ij;
        std::vector<float> w;
        rnng_host::graph_edges(graph_->h, ij, w);
        for (size_t e = 0; e < w.size(); ++e) {
            const Index u = ij[2 * e], v = ij[2 * e + 1];
            const Index cu = cell_indices[(size_t)u], cv = cell_indices[(size_t)v];
            if (cu == cv) {
                cp[(size_t)cu]->members.emplace_back(pos[(size_t)u], pos[(size_t)v], w[e]);
                continue;
            }
            const int32_t gv = ghost_index(cu, v), gu = ghost_index(cv, u);
            if (gv >= 0) cp[(size_t)cu]->ghosts.emplace_back(pos[(size_t)u], gv, w[e]);
            if (gu >= 0) cp[(size_t)cv]->ghosts.emplace_back(pos[(size_t)v], gu, w[e]);
        }
        for (Index c = 0; c < num_centers; ++c) cells[(size_t)c].serve_pairs(cp[(size_t)c]);
        return pairs;
    }

    // declared after the base's device copy it borrows, destroyed before it
    std::shared_ptr<Handle> handle_;
    mutable std::shared_ptr<rnng_host::GraphOwner> graph_;
    mutable Real built_radius = -1;
    Index num_centers = 0;
    IndexVector cell_indices;
    RealVector dist_to_centers;
};

#endif
