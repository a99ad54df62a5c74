// ggp_xengine.h -- C++ drop-in for ggp::XEngine (include/algorithm.h:49-118,
// lib/xengine.cpp:22-306) over the C ABI in ftdx.h.  Header-only; link libftdx.
//
// The parameter struct is ftdx_xengine_param, field for field the reference's
// XEngineParam (include/ggp.h:101-123) with its enum values (include/enum_ggp.h);
// a caller holding an XEngineParam copies the fields across (see to_ftdx below
// for the shape of that copy).  Errors the reference reports through errorQuda
// (a process abort) are thrown as std::runtime_error.
#pragma once

#include <stdexcept>
#include <string>

#include "ftdx.h"

namespace ggp {

class XEngine {
   public:
    explicit XEngine(const ftdx_xengine_param& param) {
        if (ftdx_xengine_create(&param, &xe_) != FTDX_OK)
            throw std::runtime_error(std::string("XEngine: ") + ftdx_last_error());
    }
    static XEngine* create(const ftdx_xengine_param& param) { return new XEngine(param); }

    XEngine(const XEngine&) = delete;
    XEngine& operator=(const XEngine&) = delete;
    ~XEngine() { ftdx_xengine_destroy(xe_); }

    // One call: every payload (lib/xengine.cpp:110-178); `stream` is a hipStream_t
    // (nullptr = default stream); returns once the stream is idle.
    void compute(void* output_data, const void* input_data, void* stream = nullptr) {
        if (ftdx_xengine_compute(xe_, output_data, input_data, stream) != FTDX_OK)
            throw std::runtime_error(std::string("XEngine::compute: ") + ftdx_last_error());
    }

    size_t in_payload_bytes() const { return query().in; }
    size_t out_payload_bytes() const { return query().out; }

   private:
    struct Sizes {
        size_t in, out;
    };
    Sizes query() const {
        Sizes s{0, 0};
        ftdx_xengine_query(xe_, &s.in, &s.out, nullptr, nullptr, nullptr);
        return s;
    }
    ftdx_xengine* xe_ = nullptr;
};

// Template for the field copy from the reference's XEngineParam (any struct with
// those member names); enum values are the same numbers.
template <typename RefParam>
inline ftdx_xengine_param to_ftdx(const RefParam& r) {
    ftdx_xengine_param p;
    ftdx_xengine_param_init(&p);
    p.struct_size = sizeof(p);
    p.in_location = (int)r.in_location;
    p.out_location = (int)r.out_location;
    p.packet_format = (int)r.packet_format;
    p.n_payload = r.n_payload;
    p.n_antennae_per_payload = r.n_antennae_per_payload;
    p.n_channels_per_payload = r.n_channels_per_payload;
    p.n_time_per_payload = r.n_time_per_payload;
    p.n_time_inner = r.n_time_inner;
    p.n_polarizations = r.n_polarizations;
    p.compute_prec = (int)r.compute_prec;
    p.output_prec = (int)r.output_prec;
    p.engine = (int)r.engine;
    p.format = (int)r.format;
    return p;
}

}  // namespace ggp
