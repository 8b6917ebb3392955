// Internal declarations of the MI355X MPFA-O Richards library.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tdyhip.h"

namespace tdyhip {

// Hexahedral interaction regions: at most 8 cells and 12 faces per vertex,
// and at most 8 transmissibility columns (cells + Dirichlet faces) per row.
constexpr int kMaxIRCells = 8;
constexpr int kMaxIRFaces = 12;
constexpr int kMaxCols = 8;

// face flags
constexpr uint8_t kFaceBoundary = 1;
constexpr uint8_t kFaceNoFlow = 2;    // Neumann boundary face: no flux row
constexpr uint8_t kFaceSeepage = 4;

// Per-vertex description of the local MPFA-O system, built on the host and
// consumed by the batched setup kernel (all indices vertex-local).
struct IRDesc {
  uint8_t ncell, nface, nint, nneu, ndir, pad[3];
  uint8_t role[kMaxIRFaces];   // 0 internal, 1 no-flow, 2 Dirichlet/seepage
  uint8_t fidx[kMaxIRFaces];   // unknown index (internal/no-flow) or Dirichlet column index
  uint8_t fup_i[kMaxIRFaces];  // local cell of the subcell on the up side (or the only one)
  uint8_t fup_l[kMaxIRFaces];  // its local face (0..2)
  uint8_t fdn_i[kMaxIRFaces];
  uint8_t fdn_l[kMaxIRFaces];
  uint8_t sck[kMaxIRCells][3]; // subcell (local cell i, face l) -> vertex face index
  int32_t row[kMaxIRFaces];    // output flux row (face*4+subface) per vertex face, -1 for no-flow
  int32_t subcell[kMaxIRCells];// global subcell id (cell*8 + corner)
};

struct HostTopology {
  int nc = 0, nf = 0, nv = 0, nbnd = 0;
  std::vector<int32_t> vcols;     // nv*kMaxCols: cell ids then nc+b, -1 padding
  std::vector<IRDesc> ir;         // nv
  std::vector<int32_t> sc_faces;  // nc*8*3 subcell face triples (oriented)
  std::vector<uint8_t> jslot;     // nv*64: position of column cell k in Jacobian row of cell i
  std::vector<int32_t> jrp, jcols, jdiag;
  std::vector<int32_t> bnd_cell;  // nbnd
  std::vector<uint8_t> face_flags;// nf
  bool over_limit = false;        // set with the error when a vertex exceeds the fixed-size limits above
};

// Builds the host topology; returns an error message or empty string.  A mesh
// that only exceeds the hexahedral kernels' fixed limits (over_limit) is
// valid: it takes the general path.
std::string build_topology(const tdyhip_mesh &m, HostTopology &t);

// Interaction regions of a general mesh (hexahedra and right prisms, padded
// with -1 in the hexahedral arrays), of any size up to kGenMaxLanes unknowns
// and columns, as variable-length records for the size-bucketed setup
// kernel.  Per vertex v, vdesc[8v..8v+7] = ncell, nface, nU, ndir, offset of
// its cells, offset of its faces, offset of its T block, T rows.
constexpr int kGenMaxLanes = 64;
struct HostGeneral {
  std::vector<int32_t> vdesc;     // nv*8
  std::vector<int32_t> csub;      // per IR cell: subcell id (cell*8 + corner)
  std::vector<int32_t> ccell;     // per IR cell: the cell (its T column)
  std::vector<int32_t> csck;      // per IR cell: vertex-face index of its 3 local faces (k0 | k1<<8 | k2<<16)
  std::vector<int32_t> fa, fb;    // per IR face: role | fidx<<8 | fup_i<<16 | fup_l<<24;
                                  //              fdn_i | fdn_l<<8 | (T row + 1)<<16
  std::vector<int32_t> frow;      // per IR face: flux row face*4+q
  std::vector<int32_t> fcol;      // per IR face: T column nc+b of a Dirichlet face, else -1
  std::vector<int32_t> bucket[3]; // vertices whose region fits 16, 32 and 64 lanes
  std::vector<int32_t> sc_faces;  // nc*8*3 subcell face triples (oriented), -1 for absent corners
  std::vector<uint8_t> face_nv;   // nf: vertices of each face (3 or 4)
  std::vector<uint8_t> face_flags;// nf
  std::vector<int32_t> bnd_cell;  // nbnd
  int64_t tsize = 0;              // T entries (sum of rows x columns)
  int max_cells = 0, max_faces = 0;
  // flux rows (face*4 + q) of the residual and Jacobian kernels
  std::vector<int32_t> rtoff;     // 4nf: offset of the row's T block row (-1: no row)
  std::vector<int32_t> rinfo;     // 4nf: ncell | ndir << 8 of the row's vertex
  std::vector<int32_t> roc, rdc;  // 4nf: offsets of the vertex's IR cells and Dirichlet columns
  std::vector<int32_t> vdcol;     // per vertex, its Dirichlet columns' boundary indices (fidx order)
  std::vector<int32_t> jrp, jcols, jdiag;  // vertex-adjacency Jacobian pattern
  int lanes[3] = {0, 0, 0};       // team width chosen for each bucket (0: empty bucket)
  float tune_ms[3][3] = {};       // tuning runs: bucket x width 16/32/64 (0: not run)
  float sched_ms[4] = {};         // tuning runs of the whole launch: order (small, large first) x chunk (1, 8)
  bool ir_desc = false;           // buckets drained largest regions first
  int ir_chunk = 1;               // workgroup steps per work item
};
std::string build_topology_general(const tdyhip_mesh &m, HostGeneral &g);
// True when every cell has 8 vertices and every face 4 (the hexahedral path).
bool mesh_is_hex(const tdyhip_mesh &m);

void set_error(const std::string &msg);

struct SolverWS;

// Per-kernel timers: HIP events around each launch on the launching stream,
// plus the launch's algorithmic bytes (each array counted once).  Launches
// inside a GMRES restart cycle carry the Arnoldi step they belong to (tag);
// once the cycle's step count is known, the launches of steps past it (which
// returned at once on the cycle's stop flag) are moved to a record of their
// own, "<kernel> (exited at once)", with no bytes, so that a kernel's record
// holds only launches that did its work.
struct KernelTimers {
  bool on = false;
  int tag = -1;  // Arnoldi step of the launches being recorded (-1: none)
  struct Pending {
    hipEvent_t e0, e1;
    int tag;
    double bytes;
  };
  struct Rec {
    int64_t launches = 0;
    double bytes = 0.0;
    double ms = 0.0;
    std::vector<Pending> pending;
  };
  std::vector<std::pair<std::string, Rec>> recs;
  std::vector<hipEvent_t> pool;
  Rec &get(const std::string &name) {
    for (auto &r : recs)
      if (r.first == name) return r.second;
    recs.emplace_back(name, Rec());
    return recs.back().second;
  }
};

}  // namespace tdyhip

struct tdyhip_fuser;

struct tdyhip_ctx {
  int nc = 0, nf = 0, nv = 0, nbnd = 0;
  int64_t jnnz = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::atomic<int> cancel{0};  // set by tdyhip_cancel from any thread
  hipEvent_t ws_event = nullptr;  // marks the stream's progress for a serving batch's fused launches
  tdyhip_fuser *fuser = nullptr;  // evaluation fuser the SNES-form evaluations go through, if any

  double vg_m, vg_alpha, x0, x1, x2, dx, gravity, pref;
  int density_type, gmatrix;
  double mualem_c[4];
  double dtime = 1.0;
  bool has_pbnd_fn = false;

  // device: mesh
  int32_t *d_face_cells = nullptr;  // nf*2
  int32_t *d_face_verts = nullptr;  // nf*4
  int32_t *d_cell_faces = nullptr;  // nc*6
  uint8_t *d_face_flags = nullptr;  // nf
  double *d_cell_vol = nullptr, *d_phi = nullptr, *d_Sr = nullptr;
  double *d_pbnd_fn = nullptr;      // nbnd
  int32_t *d_bnd_cell = nullptr;    // nbnd
  double *d_cell_z = nullptr;       // nc cell-centroid z (velocity recovery)
  double *d_face_area = nullptr;    // nf

  // device: discretisation
  int32_t *d_vcols = nullptr;       // nv*kMaxCols
  double *d_T = nullptr;            // kMaxCols * 4nf (column-major ELL over face-major rows)
  double *d_G = nullptr;            // 4nf gravity term
  uint8_t *d_jslot = nullptr;       // nv*64: Jacobian-row slot of (region cell i, column k)
  uint32_t *d_crow = nullptr;       // nc*8: a corner's three flux rows, the cell's region index, its sides
  int32_t *d_cell_verts = nullptr;  // nc*8 (hexahedra)
  int cj_xcd = 1;                   // k_cell_jac: XCD-grouped workgroup numbering (TDYHIP_CELLJAC_XCD=0: off)
  int32_t *d_jrp = nullptr, *d_jcols = nullptr;
  uint8_t *d_jdslot = nullptr;      // nc: slot of the diagonal within its row
  int jell_w = 0;                   // ELL width (longest Jacobian row)
  int32_t *d_ecols = nullptr;       // jell_w*nc slot-major ELL columns (padding = own row)

  // device: state
  double *d_rho = nullptr, *d_ukvr = nullptr, *d_drho = nullptr, *d_dukvr = nullptr;
  double *d_dM = nullptr, *d_d2M = nullptr, *d_acc = nullptr;
  double *d_Pbnd = nullptr, *d_ukvr_bnd = nullptr, *d_rho_bnd = nullptr;
  double *d_flux = nullptr;         // 4nf
  double *d_accum_prev = nullptr;   // nc
  double *d_source = nullptr;       // nc source_sink (tdyhip_set_source_sink)
  bool has_source = false;

  // host copies used by diagnostics
  std::vector<int32_t> h_jrp, h_jcols;
  std::vector<int32_t> h_vcols;
  std::vector<int32_t> h_face_verts;
  std::vector<uint8_t> h_face_flags;
  std::vector<void *> allocations;
  int64_t dev_bytes = 0;           // bytes held in `allocations`
  tdyhip::SolverWS *ws = nullptr;  // Newton/Krylov workspace, allocated on first solve
  std::vector<int32_t> pc_part;    // caller-given block-Jacobi partition (empty: breadth-first)
  int pc_part_version = 0;
  tdyhip::KernelTimers kt;
  // Set by the Newton solver around its own residual / Jacobian launches (all
  // on ctx->stream): device-resident dtime and shift, and a device stop flag
  // that makes the launches return at once.  Null for the ABI callbacks.
  const double *launch_dyn = nullptr;
  const int *launch_stop = nullptr;
  double *launch_flux = nullptr;  // flux-row scratch in place of d_flux (the allocation-free residual)
  // algorithmic-byte bookkeeping
  int64_t trans_nnz = 0;           // transmissibility entries (cells + Dirichlet columns)
  int64_t active_rows = 0;         // flux rows of faces that carry flux
  // general (prism) meshes: transmissibility setup only, T as per-vertex blocks
  bool general = false;
  double *d_Tgen = nullptr;        // sum over vertices of rows x columns, row-major blocks
  double *d_Dgen = nullptr;        // flux-row derivatives, same blocks (cell columns)
  int32_t *d_rtoff = nullptr, *d_rinfo = nullptr, *d_roc = nullptr, *d_rdc = nullptr;  // 4nf each
  int32_t *d_ccell = nullptr, *d_vdcol = nullptr, *d_jdiag = nullptr;
  uint8_t *d_gslot = nullptr;      // 2 per block entry: Jacobian slot in the up / down cell's row
  int32_t *d_vblk = nullptr;       // nv*7 vertex blocks for the row kernel (nullptr: row per thread)
  int32_t *d_brow = nullptr;       // flux row of each block row
  tdyhip::HostGeneral *hgen = nullptr;
  double setup_ms[3] = {0, 0, 0};  // setup device times: G matrices, local systems, gravity term
  double ls_flops = 0.0, ls_bytes = 0.0;  // hexahedral local systems: algorithmic FP64 flops and bytes
  int ir_max_cells = 0, ir_max_faces = 0;
};
