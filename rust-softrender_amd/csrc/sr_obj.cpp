// sr_obj.cpp — Wavefront OBJ ingestion (sr_obj_parse, sr_mesh_load_obj).
//
// The reference's examples load meshes with the `tobj` crate 0.1.3
// (`tobj::load_obj`, examples/suzanne.rs:23-47, full_example/src/main.rs);
// the crate is not vendored, so this restates its published behaviour:
//   * `v`, `vt`, `vn` lines append positions, texture coordinates and
//     normals; `f` lines list corners `v`, `v/vt`, `v//vn` or `v/vt/vn`
//     (1-based; negative = relative to the current count);
//   * an `o` line starts a new model once faces have been read; other
//     statements (`g`, `s`, `usemtl`, `mtllib`, comments) are skipped;
//   * polygons become a triangle fan from their first corner:
//     (c0, c1, c2), (c0, c2, c3), ...;
//   * each distinct (v, vt, vn) corner becomes one output vertex, numbered
//     in order of first use; the index list refers to those vertices.
// The mesh then feeds `Mesh { vertices, indices }` (src/mesh.rs:12-64)
// exactly as suzanne.rs builds it (position + normal per vertex).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "sr_runtime.hpp"

namespace {

struct ObjModel {
    std::vector<float> pos, nrm, uv;
    std::vector<uint32_t> idx;
    bool has_n = false, has_t = false;
};

typedef std::tuple<long, long, long> Corner;  // (v, vt, vn), 0-based, -1 = absent

long resolve(long i, size_t count) {
    if (i > 0) return i - 1;
    if (i < 0) return (long)count + i;
    return -2;  // 0 is not a valid OBJ index
}

bool parse_corner(const char* tok, size_t nv, size_t nt, size_t nn, Corner& c) {
    long v = 0, t = 0, n = 0;
    char* e = nullptr;
    v = strtol(tok, &e, 10);
    if (e == tok) return false;
    long rv = resolve(v, nv), rt = -1, rn = -1;
    if (*e == '/') {
        const char* p = e + 1;
        if (*p != '/') {
            t = strtol(p, &e, 10);
            if (e == p) return false;
            rt = resolve(t, nt);
        } else {
            e = (char*)p;
        }
        if (*e == '/') {
            p = e + 1;
            n = strtol(p, &e, 10);
            if (e == p) return false;
            rn = resolve(n, nn);
        }
    }
    if (rv < 0 || rv >= (long)nv || (t && (rt < 0 || rt >= (long)nt)) || (n && (rn < 0 || rn >= (long)nn)))
        return false;
    c = Corner(rv, rt, rn);
    return true;
}

int parse_obj(const char* path, std::vector<ObjModel>& models, std::string& err) {
    FILE* f = fopen(path, "rb");
    if (!f) {
        err = std::string("obj: cannot open ") + path;
        return SR_ERR_INVALID_ARGUMENT;
    }
    std::vector<float> P, T, N;
    std::vector<std::vector<Corner>> faces;
    auto flush = [&]() {
        if (faces.empty()) return;
        ObjModel m;
        std::map<Corner, uint32_t> seen;
        auto add = [&](const Corner& c) {
            auto it = seen.find(c);
            if (it != seen.end()) {
                m.idx.push_back(it->second);
                return;
            }
            const uint32_t id = (uint32_t)seen.size();
            seen.emplace(c, id);
            const long v = std::get<0>(c), t = std::get<1>(c), n = std::get<2>(c);
            m.pos.insert(m.pos.end(), {P[3 * v], P[3 * v + 1], P[3 * v + 2]});
            if (t >= 0) {
                m.uv.insert(m.uv.end(), {T[2 * t], T[2 * t + 1]});
                m.has_t = true;
            }
            if (n >= 0) {
                m.nrm.insert(m.nrm.end(), {N[3 * n], N[3 * n + 1], N[3 * n + 2]});
                m.has_n = true;
            }
            m.idx.push_back(id);
        };
        for (const auto& fc : faces)
            for (size_t k = 2; k < fc.size(); k++) {
                add(fc[0]);
                add(fc[k - 1]);
                add(fc[k]);
            }
        models.push_back(std::move(m));
        faces.clear();
    };
    char line[4096];
    int rc = SR_OK;
    while (fgets(line, sizeof(line), f)) {
        char* s = line;
        while (*s == ' ' || *s == '\t') s++;
        if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
            float x = 0, y = 0, z = 0;
            if (sscanf(s + 2, "%f %f %f", &x, &y, &z) != 3) { rc = SR_ERR_INVALID_ARGUMENT; break; }
            P.insert(P.end(), {x, y, z});
        } else if (s[0] == 'v' && s[1] == 't' && (s[2] == ' ' || s[2] == '\t')) {
            float u = 0, v = 0;
            if (sscanf(s + 3, "%f %f", &u, &v) != 2) { rc = SR_ERR_INVALID_ARGUMENT; break; }
            T.insert(T.end(), {u, v});
        } else if (s[0] == 'v' && s[1] == 'n' && (s[2] == ' ' || s[2] == '\t')) {
            float x = 0, y = 0, z = 0;
            if (sscanf(s + 3, "%f %f %f", &x, &y, &z) != 3) { rc = SR_ERR_INVALID_ARGUMENT; break; }
            N.insert(N.end(), {x, y, z});
        } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
            std::vector<Corner> fc;
            char* save = nullptr;
            for (char* tok = strtok_r(s + 2, " \t\r\n", &save); tok; tok = strtok_r(nullptr, " \t\r\n", &save)) {
                Corner c;
                if (!parse_corner(tok, P.size() / 3, T.size() / 2, N.size() / 3, c)) { rc = SR_ERR_INVALID_ARGUMENT; break; }
                fc.push_back(c);
            }
            if (rc) break;
            if (fc.size() < 3) { rc = SR_ERR_INVALID_ARGUMENT; break; }
            faces.push_back(std::move(fc));
        } else if (s[0] == 'o' && (s[1] == ' ' || s[1] == '\t' || s[1] == '\n' || s[1] == '\r')) {
            flush();
        }
    }
    fclose(f);
    if (rc) {
        err = std::string("obj: malformed statement in ") + path;
        return rc;
    }
    flush();
    for (const ObjModel& m : models)
        if ((m.has_n && m.nrm.size() / 3 != m.pos.size() / 3) || (m.has_t && m.uv.size() / 2 != m.pos.size() / 3)) {
            err = "obj: corners mix with and without normals or texture coordinates";
            return SR_ERR_UNSUPPORTED;
        }
    return SR_OK;
}

}  // namespace

extern "C" {

int sr_obj_parse(const char* path, uint32_t model, sr_obj_info* info, float* positions, float* normals,
                 float* uvs, uint32_t* indices) {
    if (!path || !info) return SR_ERR_INVALID_ARGUMENT;
    std::vector<ObjModel> models;
    std::string err;
    int rc = parse_obj(path, models, err);
    if (rc) return rc;
    memset(info, 0, sizeof(*info));
    info->models = (uint32_t)models.size();
    if (model >= models.size()) return models.empty() ? SR_OK : SR_ERR_INVALID_ARGUMENT;
    const ObjModel& m = models[model];
    info->vertices = (uint32_t)(m.pos.size() / 3);
    info->indices = (uint32_t)m.idx.size();
    info->has_normals = m.has_n;
    info->has_uvs = m.has_t;
    if (positions) memcpy(positions, m.pos.data(), sizeof(float) * m.pos.size());
    if (normals && m.has_n) memcpy(normals, m.nrm.data(), sizeof(float) * m.nrm.size());
    if (uvs && m.has_t) memcpy(uvs, m.uv.data(), sizeof(float) * m.uv.size());
    if (indices) memcpy(indices, m.idx.data(), sizeof(uint32_t) * m.idx.size());
    return SR_OK;
}

int sr_mesh_load_obj(sr_context ctx, const char* path, uint32_t model, sr_mesh* out) {
    if (!ctx || !path || !out) return SR_ERR_INVALID_ARGUMENT;
    std::vector<ObjModel> models;
    std::string err;
    int rc = parse_obj(path, models, err);
    if (rc) return sr_rt::fail(ctx, rc, err);
    if (model >= models.size()) return sr_rt::fail(ctx, SR_ERR_INVALID_ARGUMENT, "obj: no such model");
    const ObjModel& m = models[model];
    const size_t nv = m.pos.size() / 3;
    std::vector<float> soa(8 * nv, 0.0f);
    for (size_t i = 0; i < nv; i++) {
        for (int c = 0; c < 3; c++) soa[c * nv + i] = m.pos[3 * i + c];
        if (m.has_n)
            for (int c = 0; c < 3; c++) soa[(3 + c) * nv + i] = m.nrm[3 * i + c];
        if (m.has_t)
            for (int c = 0; c < 2; c++) soa[(6 + c) * nv + i] = m.uv[2 * i + c];
    }
    sr_vertex_soa v;
    v.px = &soa[0]; v.py = &soa[nv]; v.pz = &soa[2 * nv];
    v.nx = m.has_n ? &soa[3 * nv] : nullptr;
    v.ny = m.has_n ? &soa[4 * nv] : nullptr;
    v.nz = m.has_n ? &soa[5 * nv] : nullptr;
    v.u = m.has_t ? &soa[6 * nv] : nullptr;
    v.v = m.has_t ? &soa[7 * nv] : nullptr;
    return sr_mesh_upload(ctx, &v, (uint32_t)nv, m.idx.data(), (uint32_t)m.idx.size(), SR_PRIM_TRIANGLES, out);
}

}  // extern "C"
