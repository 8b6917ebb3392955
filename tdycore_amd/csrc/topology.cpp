// Host-side topology of the MPFA-O interaction regions (integer work only).
//
// Plays the part of TDyMeshCreateFromPlex's vertex/subcell bookkeeping
// (src/fv/share/tdymeshplex.c:1531-1713 SetupUpwindFacesForSubcell,
// 2829-2880 SetupSubcells, 236-275 FindFaceIDsOfACellCommonToAVertex,
// 1384-1412 UpdateFaceOrientationAroundAVertex) and of the vertex-adjacency
// Jacobian pattern (src/tdydiscretization.c:156-181).  The floating-point
// work (G matrices, local solves, gravity) runs on the device.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "tdyhip_internal.h"

namespace tdyhip {

static int face_pos(const int32_t *fv, int f, int v) {
  for (int q = 0; q < 4; q++)
    if (fv[f * 4 + q] == v) return q;
  return -1;
}

std::string build_topology(const tdyhip_mesh &m, HostTopology &t) {
  const int nc = m.nc, nf = m.nf, nv = m.nv, nbnd = m.nbnd;
  t.nc = nc; t.nf = nf; t.nv = nv; t.nbnd = nbnd;
  const int32_t *cv = m.cell_verts, *cf = m.cell_faces, *fv = m.face_verts, *fc = m.face_cells;

  t.face_flags.assign(nf, 0);
  t.bnd_cell.assign(nbnd, -1);
  for (int f = 0; f < nf; f++) {
    int up = fc[2 * f], dn = fc[2 * f + 1];
    if (up < 0 && dn < 0) return "face with two boundary sides";
    if (up < 0 || dn < 0) {
      if (up >= 0) return "boundary face must be stored as (-(b+1), cell)";
      int b = -up - 1;
      if (b >= nbnd) return "boundary index out of range";
      t.bnd_cell[b] = dn;
      uint8_t fl = kFaceBoundary;
      int bc = m.face_bc[f];
      if (bc == TDYHIP_BC_NEUMANN) fl |= kFaceNoFlow;
      else if (bc == TDYHIP_BC_SEEPAGE) fl |= kFaceSeepage;
      else if (bc != TDYHIP_BC_DIRICHLET) return "boundary face without a boundary type";
      t.face_flags[f] = fl;
    }
  }
  for (int b = 0; b < nbnd; b++)
    if (t.bnd_cell[b] < 0) return "boundary index without a face";

  // subcell face triples, oriented (normal of the face-centroid triangle toward the cell centroid)
  t.sc_faces.assign((size_t)nc * 24, -1);
  for (int c = 0; c < nc; c++) {
    for (int s = 0; s < 8; s++) {
      int v = cv[c * 8 + s];
      int fid[3], n = 0;
      for (int lf = 0; lf < 6; lf++) {
        int f = cf[c * 6 + lf];
        if (face_pos(fv, f, v) >= 0) {
          if (n < 3) fid[n] = f;
          n++;
        }
      }
      if (n != 3) return "hexahedral subcell without exactly 3 faces at its vertex";
      double a[3], b[3], cvec[3];
      for (int d = 0; d < 3; d++) {
        a[d] = m.face_cen[3 * fid[1] + d] - m.face_cen[3 * fid[0] + d];
        b[d] = m.face_cen[3 * fid[2] + d] - m.face_cen[3 * fid[0] + d];
        cvec[d] = m.cell_cen[3 * c + d] - m.vert_xyz[3 * v + d];
      }
      double axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
      if (axb[0] * cvec[0] + axb[1] * cvec[1] + axb[2] * cvec[2] < 0) std::swap(fid[1], fid[2]);
      for (int k = 0; k < 3; k++) t.sc_faces[(size_t)(c * 8 + s) * 3 + k] = fid[k];
    }
  }

  // vertex -> cells (cell order), vertex -> faces (internal then boundary, face order)
  std::vector<int32_t> vc_n(nv, 0), vf_n(nv, 0);
  std::vector<int32_t> vcell((size_t)nv * kMaxIRCells, -1), vsub((size_t)nv * kMaxIRCells, -1);
  std::vector<int32_t> vface((size_t)nv * kMaxIRFaces, -1), vsubf((size_t)nv * kMaxIRFaces, -1);
  for (int c = 0; c < nc; c++)
    for (int s = 0; s < 8; s++) {
      int v = cv[c * 8 + s];
      if (v < 0 || v >= nv) return "cell vertex out of range";
      if (vc_n[v] >= kMaxIRCells) { t.over_limit = true; return "more than 8 cells at a vertex"; }
      vcell[(size_t)v * kMaxIRCells + vc_n[v]] = c;
      vsub[(size_t)v * kMaxIRCells + vc_n[v]] = s;
      vc_n[v]++;
    }
  for (int pass = 0; pass < 2; pass++)
    for (int f = 0; f < nf; f++) {
      if (((t.face_flags[f] & kFaceBoundary) != 0) != (pass == 1)) continue;
      for (int q = 0; q < 4; q++) {
        int v = fv[f * 4 + q];
        if (vf_n[v] >= kMaxIRFaces) { t.over_limit = true; return "more than 12 faces at a vertex"; }
        vface[(size_t)v * kMaxIRFaces + vf_n[v]] = f;
        vsubf[(size_t)v * kMaxIRFaces + vf_n[v]] = q;
        vf_n[v]++;
      }
    }

  t.ir.assign(nv, IRDesc{});
  t.vcols.assign((size_t)nv * kMaxCols, -1);
  for (int v = 0; v < nv; v++) {
    IRDesc &d = t.ir[v];
    std::memset(&d, 0, sizeof(d));
    for (int k = 0; k < kMaxIRFaces; k++) d.row[k] = -1;
    int ncell = vc_n[v], nface = vf_n[v];
    d.ncell = (uint8_t)ncell;
    d.nface = (uint8_t)nface;
    int nint = 0, nneu = 0, ndir = 0;
    for (int k = 0; k < nface; k++)
      if (!(t.face_flags[vface[(size_t)v * kMaxIRFaces + k]] & kFaceBoundary)) nint++;
    for (int k = 0; k < nface; k++) {
      int f = vface[(size_t)v * kMaxIRFaces + k];
      uint8_t fl = t.face_flags[f];
      if (!(fl & kFaceBoundary)) { d.role[k] = 0; d.fidx[k] = (uint8_t)k; }
      else if (fl & kFaceNoFlow) { d.role[k] = 1; d.fidx[k] = (uint8_t)(nint + nneu); nneu++; }
      else { d.role[k] = 2; d.fidx[k] = (uint8_t)ndir; ndir++; }
      d.fup_i[k] = d.fdn_i[k] = 255;
    }
    d.nint = (uint8_t)nint; d.nneu = (uint8_t)nneu; d.ndir = (uint8_t)ndir;
    if (ncell + ndir > kMaxCols) { t.over_limit = true; return "more than 8 transmissibility columns at a vertex"; }
    // columns: cells, then Dirichlet faces (boundary index)
    for (int i = 0; i < ncell; i++) {
      t.vcols[(size_t)v * kMaxCols + i] = vcell[(size_t)v * kMaxIRCells + i];
      d.subcell[i] = vcell[(size_t)v * kMaxIRCells + i] * 8 + vsub[(size_t)v * kMaxIRCells + i];
    }
    for (int k = 0; k < nface; k++)
      if (d.role[k] == 2) {
        int f = vface[(size_t)v * kMaxIRFaces + k];
        t.vcols[(size_t)v * kMaxCols + ncell + d.fidx[k]] = nc + (-fc[2 * f] - 1);
      }
    // subcell faces -> vertex face index; up/down subcell of each face
    for (int i = 0; i < ncell; i++) {
      int c = vcell[(size_t)v * kMaxIRCells + i], s = vsub[(size_t)v * kMaxIRCells + i];
      for (int l = 0; l < 3; l++) {
        int f = t.sc_faces[(size_t)(c * 8 + s) * 3 + l];
        int k = 0;
        while (k < nface && vface[(size_t)v * kMaxIRFaces + k] != f) k++;
        if (k == nface) return "subcell face not at its vertex";
        d.sck[i][l] = (uint8_t)k;
        if (fc[2 * f] == c || (t.face_flags[f] & kFaceBoundary)) { d.fup_i[k] = (uint8_t)i; d.fup_l[k] = (uint8_t)l; }
        else { d.fdn_i[k] = (uint8_t)i; d.fdn_l[k] = (uint8_t)l; }
      }
    }
    for (int k = 0; k < nface; k++) {
      if (d.fup_i[k] == 255 || (d.role[k] == 0 && d.fdn_i[k] == 255)) return "face without its subcells";
      int f = vface[(size_t)v * kMaxIRFaces + k], q = vsubf[(size_t)v * kMaxIRFaces + k];
      if (d.role[k] != 1) d.row[k] = f * 4 + q;
    }
  }

  // Jacobian pattern: cells sharing a vertex, sorted
  t.jrp.assign(nc + 1, 0);
  std::vector<int32_t> mark(nc, -1), tmp;
  tmp.reserve(64);
  for (int pass = 0; pass < 2; pass++) {
    for (int c = 0; c < nc; c++) {
      tmp.clear();
      for (int s = 0; s < 8; s++) {
        int v = cv[c * 8 + s];
        for (int i = 0; i < vc_n[v]; i++) {
          int cc = vcell[(size_t)v * kMaxIRCells + i];
          if (mark[cc] != c) { mark[cc] = c; tmp.push_back(cc); }
        }
      }
      if (pass == 0) t.jrp[c + 1] = (int32_t)tmp.size();
      else {
        std::sort(tmp.begin(), tmp.end());
        std::copy(tmp.begin(), tmp.end(), t.jcols.begin() + t.jrp[c]);
      }
    }
    if (pass == 0) {
      for (int c = 0; c < nc; c++) t.jrp[c + 1] += t.jrp[c];
      t.jcols.assign(t.jrp[nc], 0);
      std::fill(mark.begin(), mark.end(), -1);
    }
  }
  t.jdiag.assign(nc, -1);
  for (int c = 0; c < nc; c++)
    for (int k = t.jrp[c]; k < t.jrp[c + 1]; k++)
      if (t.jcols[k] == c) t.jdiag[c] = k;
  // slot of (row cell i, column cell k) within row i, per vertex
  t.jslot.assign((size_t)nv * 64, 255);
  for (int v = 0; v < nv; v++) {
    int ncell = vc_n[v];
    for (int i = 0; i < ncell; i++) {
      int ci = vcell[(size_t)v * kMaxIRCells + i];
      const int32_t *row = t.jcols.data() + t.jrp[ci];
      int len = t.jrp[ci + 1] - t.jrp[ci];
      if (len > 255) return "Jacobian row longer than 255";
      for (int k = 0; k < ncell; k++) {
        int ck = vcell[(size_t)v * kMaxIRCells + k];
        int pos = (int)(std::lower_bound(row, row + len, ck) - row);
        t.jslot[(size_t)v * 64 + i * 8 + k] = (uint8_t)pos;
      }
    }
  }
  return std::string();
}

bool mesh_is_hex(const tdyhip_mesh &m) {
  for (int64_t i = 0; i < (int64_t)m.nc * 8; i++)
    if (m.cell_verts[i] < 0) return false;
  for (int64_t i = 0; i < (int64_t)m.nc * 6; i++)
    if (m.cell_faces[i] < 0) return false;
  for (int64_t i = 0; i < (int64_t)m.nf * 4; i++)
    if (m.face_verts[i] < 0) return false;
  return true;
}

// The same bookkeeping as build_topology, for cells and faces of any vertex
// count (the reference's general path: ComputeTransmissibilityMatrix over
// vertices of any degree, tdympfao.c:1801-1860, with the local systems of
// 1126-1513 sized by the region).
std::string build_topology_general(const tdyhip_mesh &m, HostGeneral &g) {
  const int nc = m.nc, nf = m.nf, nv = m.nv, nbnd = m.nbnd;
  const int32_t *cv = m.cell_verts, *cf = m.cell_faces, *fv = m.face_verts, *fc = m.face_cells;
  g.face_flags.assign(nf, 0);
  g.face_nv.assign(nf, 0);
  g.bnd_cell.assign(nbnd, -1);
  for (int f = 0; f < nf; f++) {
    int n = 0;
    for (int q = 0; q < 4; q++) {
      const int v = fv[f * 4 + q];
      if (v >= nv) return "face vertex out of range";
      if (v >= 0) {
        if (q != n) return "face vertices must come first, then -1 padding";
        n++;
      }
    }
    if (n < 3) return "face with fewer than 3 vertices";
    g.face_nv[f] = (uint8_t)n;
    int up = fc[2 * f], dn = fc[2 * f + 1];
    if (up < 0 && dn < 0) return "face with two boundary sides";
    if (up < 0 || dn < 0) {
      if (up >= 0) return "boundary face must be stored as (-(b+1), cell)";
      int b = -up - 1;
      if (b >= nbnd) return "boundary index out of range";
      g.bnd_cell[b] = dn;
      uint8_t fl = kFaceBoundary;
      int bc = m.face_bc[f];
      if (bc == TDYHIP_BC_NEUMANN) fl |= kFaceNoFlow;
      else if (bc == TDYHIP_BC_SEEPAGE) fl |= kFaceSeepage;
      else if (bc != TDYHIP_BC_DIRICHLET) return "boundary face without a boundary type";
      g.face_flags[f] = fl;
    }
  }
  for (int b = 0; b < nbnd; b++)
    if (g.bnd_cell[b] < 0) return "boundary index without a face";

  auto fpos = [&](int f, int v) {
    for (int q = 0; q < g.face_nv[f]; q++)
      if (fv[f * 4 + q] == v) return q;
    return -1;
  };
  g.sc_faces.assign((size_t)nc * 24, -1);
  std::vector<int32_t> vc_n(nv + 1, 0), vf_n(nv + 1, 0);
  for (int c = 0; c < nc; c++) {
    for (int s = 0; s < 8; s++) {
      const int v = cv[c * 8 + s];
      if (v < 0) continue;
      if (v >= nv) return "cell vertex out of range";
      vc_n[v + 1]++;
      int fid[3], n = 0;
      for (int lf = 0; lf < 6; lf++) {
        const int f = cf[c * 6 + lf];
        if (f < 0) continue;
        if (f >= nf) return "cell face out of range";
        if (fpos(f, v) >= 0) {
          if (n < 3) fid[n] = f;
          n++;
        }
      }
      if (n != 3) return "subcell without exactly 3 faces at its vertex";
      double a[3], b[3], cvec[3];
      for (int d = 0; d < 3; d++) {
        a[d] = m.face_cen[3 * fid[1] + d] - m.face_cen[3 * fid[0] + d];
        b[d] = m.face_cen[3 * fid[2] + d] - m.face_cen[3 * fid[0] + d];
        cvec[d] = m.cell_cen[3 * c + d] - m.vert_xyz[3 * v + d];
      }
      double axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
      if (axb[0] * cvec[0] + axb[1] * cvec[1] + axb[2] * cvec[2] < 0) std::swap(fid[1], fid[2]);
      for (int k = 0; k < 3; k++) g.sc_faces[(size_t)(c * 8 + s) * 3 + k] = fid[k];
    }
  }
  for (int f = 0; f < nf; f++)
    for (int q = 0; q < g.face_nv[f]; q++) vf_n[fv[f * 4 + q] + 1]++;
  for (int v = 0; v < nv; v++) { vc_n[v + 1] += vc_n[v]; vf_n[v + 1] += vf_n[v]; }
  // vertex -> (cell, corner) in cell order; vertex -> (face, q), internal faces first
  std::vector<int32_t> vcell(vc_n[nv]), vsub(vc_n[nv]), vface(vf_n[nv]), vq(vf_n[nv]);
  {
    std::vector<int32_t> fill(nv, 0);
    for (int c = 0; c < nc; c++)
      for (int s = 0; s < 8; s++) {
        const int v = cv[c * 8 + s];
        if (v < 0) continue;
        const int k = vc_n[v] + fill[v]++;
        vcell[k] = c; vsub[k] = s;
      }
    std::fill(fill.begin(), fill.end(), 0);
    for (int pass = 0; pass < 2; pass++)
      for (int f = 0; f < nf; f++) {
        if (((g.face_flags[f] & kFaceBoundary) != 0) != (pass == 1)) continue;
        for (int q = 0; q < g.face_nv[f]; q++) {
          const int v = fv[f * 4 + q];
          const int k = vf_n[v] + fill[v]++;
          vface[k] = f; vq[k] = q;
        }
      }
  }
  g.vdesc.assign((size_t)nv * 8, 0);
  g.rtoff.assign((size_t)4 * nf, -1);
  g.rinfo.assign((size_t)4 * nf, 0);
  g.roc.assign((size_t)4 * nf, 0);
  g.rdc.assign((size_t)4 * nf, 0);
  g.vdcol.clear();
  g.csub.assign(vc_n[nv], -1); g.ccell.assign(vc_n[nv], -1); g.csck.assign(vc_n[nv], 0);
  g.fa.assign(vf_n[nv], 0); g.fb.assign(vf_n[nv], 0); g.frow.assign(vf_n[nv], -1); g.fcol.assign(vf_n[nv], -1);
  for (int b = 0; b < 3; b++) g.bucket[b].clear();
  int64_t toff = 0;
  for (int v = 0; v < nv; v++) {
    const int oc = vc_n[v], of = vf_n[v];
    const int ncell = vc_n[v + 1] - oc, nface = vf_n[v + 1] - of;
    g.max_cells = std::max(g.max_cells, ncell);
    g.max_faces = std::max(g.max_faces, nface);
    int nint = 0, nneu = 0, ndir = 0;
    for (int k = 0; k < nface; k++)
      if (!(g.face_flags[vface[of + k]] & kFaceBoundary)) nint++;
    std::vector<int> role(nface), fidx(nface), fup_i(nface, -1), fup_l(nface, 0), fdn_i(nface, -1), fdn_l(nface, 0);
    for (int k = 0; k < nface; k++) {
      const int f = vface[of + k];
      const uint8_t fl = g.face_flags[f];
      if (!(fl & kFaceBoundary)) { role[k] = 0; fidx[k] = k; }
      else if (fl & kFaceNoFlow) { role[k] = 1; fidx[k] = nint + nneu; nneu++; }
      else { role[k] = 2; fidx[k] = ndir; ndir++; g.fcol[of + k] = nc + (-fc[2 * f] - 1); }
    }
    const int nU = nint + nneu, nB = ncell + ndir;
    const int need = std::max(nU, nB);
    if (need > kGenMaxLanes || nface > 255) return "interaction region larger than 64 unknowns or columns";
    for (int i = 0; i < ncell; i++) {
      const int c = vcell[oc + i], s = vsub[oc + i];
      g.csub[oc + i] = c * 8 + s;
      g.ccell[oc + i] = c;
      int packed = 0;
      for (int l = 0; l < 3; l++) {
        const int f = g.sc_faces[(size_t)(c * 8 + s) * 3 + l];
        int k = 0;
        while (k < nface && vface[of + k] != f) k++;
        if (k == nface) return "subcell face not at its vertex";
        packed |= k << (8 * l);
        if (fc[2 * f] == c || (g.face_flags[f] & kFaceBoundary)) { fup_i[k] = i; fup_l[k] = l; }
        else { fdn_i[k] = i; fdn_l[k] = l; }
      }
      g.csck[oc + i] = packed;
    }
    // T rows: internal faces, then Dirichlet faces, in vertex face order
    std::vector<int> trow(nface, -1);
    int nrow = 0;
    for (int pass = 0; pass < 2; pass++)
      for (int k = 0; k < nface; k++)
        if ((pass == 0 && role[k] == 0) || (pass == 1 && role[k] == 2)) trow[k] = nrow++;
    for (int k = 0; k < nface; k++) {
      if (fup_i[k] < 0 || (role[k] == 0 && fdn_i[k] < 0)) return "face without its subcells";
      g.fa[of + k] = role[k] | (fidx[k] << 8) | (fup_i[k] << 16) | (fup_l[k] << 24);
      g.fb[of + k] = std::max(fdn_i[k], 0) | (fdn_l[k] << 8) | ((trow[k] + 1) << 16);
      g.frow[of + k] = vface[of + k] * 4 + vq[of + k];
    }
    int32_t *d = g.vdesc.data() + (size_t)v * 8;
    d[0] = ncell; d[1] = nface; d[2] = nU; d[3] = ndir; d[4] = oc; d[5] = of; d[6] = (int32_t)toff; d[7] = nrow;
    // flux rows of this vertex: where their T rows and columns are
    const int od = (int)g.vdcol.size();
    g.vdcol.resize(od + ndir);
    for (int k = 0; k < nface; k++)
      if (role[k] == 2) g.vdcol[od + fidx[k]] = -fc[2 * vface[of + k]] - 1;
    for (int k = 0; k < nface; k++) {
      if (trow[k] < 0) continue;
      const int r = g.frow[of + k];
      g.rtoff[r] = (int32_t)(toff + (int64_t)trow[k] * (ncell + ndir));
      g.rinfo[r] = ncell | (ndir << 8);
      g.roc[r] = oc;
      g.rdc[r] = od;
    }
    if (toff + (int64_t)nrow * nB > INT32_MAX) return "transmissibility blocks exceed 2^31 entries";
    toff += (int64_t)nrow * nB;
    g.bucket[need <= 16 ? 0 : need <= 32 ? 1 : 2].push_back(v);
  }
  g.tsize = toff;
  // Jacobian pattern: cells sharing a vertex, sorted (src/tdydiscretization.c:156-181)
  g.jrp.assign(nc + 1, 0);
  std::vector<int32_t> mark(nc, -1), tmp;
  for (int pass = 0; pass < 2; pass++) {
    for (int c = 0; c < nc; c++) {
      tmp.clear();
      for (int s = 0; s < 8; s++) {
        const int v = cv[c * 8 + s];
        if (v < 0) continue;
        for (int i = vc_n[v]; i < vc_n[v + 1]; i++)
          if (mark[vcell[i]] != c) { mark[vcell[i]] = c; tmp.push_back(vcell[i]); }
      }
      if (pass == 0) g.jrp[c + 1] = (int32_t)tmp.size();
      else {
        std::sort(tmp.begin(), tmp.end());
        std::copy(tmp.begin(), tmp.end(), g.jcols.begin() + g.jrp[c]);
      }
    }
    if (pass == 0) {
      for (int c = 0; c < nc; c++) g.jrp[c + 1] += g.jrp[c];
      g.jcols.assign(g.jrp[nc], 0);
      std::fill(mark.begin(), mark.end(), -1);
    }
  }
  g.jdiag.assign(nc, -1);
  for (int c = 0; c < nc; c++)
    for (int k = g.jrp[c]; k < g.jrp[c + 1]; k++)
      if (g.jcols[k] == c) g.jdiag[c] = k - g.jrp[c];
  return std::string();
}

}  // namespace tdyhip
