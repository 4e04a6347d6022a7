// hf_faces.hpp -- the reference's HF contraction faces, and the adapter that puts the MI355X engine
// behind them.
//
//   Orbital_HF_IBS<T>       src/BasisSet/Orbital_HF_IBS.C:33-54   (+= semantics, per irrep pair)
//   WholeSystemFock_IBS<T>  src/BasisSet/Orbital_HF_IBS.C:75-87   (one whole-system build, then slice)
//
// GpuHF_IBS implements both for the (single-irrep) molecular Cartesian basis.  The caller -- the
// reference's Vee / Vxc terms through Composite_HFSystem::AccumulateDirectAll/ExchangeAll
// (ChargeDensity/Imp/CompositeCD.C:78-118) -- asks for J and then K of the same density; the
// adapter builds both in ONE engine call and serves the second request from a memo keyed by the
// density's contents (the fused J/K build the reference's ContractAll would otherwise do twice).
// Engine status codes become std::runtime_error, as Orbital_ERI4_IBS.C:19-32 does for misuse.
#pragma once

#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "linalg.hpp"
#include "qcjk.h"

namespace qchem_amd {

template <class T>
class Orbital_HF_IBS {
public:
    virtual ~Orbital_HF_IBS() = default;
    // S_ab += sum_cd (ab|cd) D_cd
    virtual void AccumulateDirect(Mat& Sab, const Mat& Dcd, const Orbital_HF_IBS<T>* bs_cd) const = 0;
    // S_ab += sum_cd (ad|cb) D_cd
    virtual void AccumulateExchange(Mat& Sab, const Mat& Dcd, const Orbital_HF_IBS<T>* bs_cd) const = 0;
    virtual void AccumulateDirectBoth(Mat& Ji, Mat& Jj, const Mat& Di, const Mat& Dj, const Orbital_HF_IBS<T>* cd) const = 0;
    virtual void AccumulateExchangeBoth(Mat& Ki, Mat& Kj, const Mat& Di, const Mat& Dj, const Orbital_HF_IBS<T>* cd) const = 0;
};

template <class T>
class WholeSystemFock_IBS {
public:
    virtual ~WholeSystemFock_IBS() = default;
    virtual size_t AODimension() const = 0;
    virtual void AddAODensity(Mat& Dao, const Mat& D) const = 0;
    virtual Mat MakeAOFock(const Mat& Dao, bool exchange) const = 0;
    virtual void SliceAOFock(Mat& Fab, const Mat& Fao) const = 0;
};

struct EngineError : std::runtime_error {
    int code;
    EngineError(const std::string& what, int c) : std::runtime_error(what + ": " + qcjk_strerror(c)), code(c) {}
};

class GpuHF_IBS : public Orbital_HF_IBS<double>, public WholeSystemFock_IBS<double> {
public:
    GpuHF_IBS(const qcjk_basis& b, const qcjk_opts& o, unsigned build_flags = QCJK_J | QCJK_K)
        : flags_(build_flags)
    {
        const int rc = qcjk_create(&b, &o, &h_);
        if (rc != QCJK_OK) throw EngineError("qcjk_create", rc);
        n_ = qcjk_nbf(h_);
    }
    ~GpuHF_IBS() override { qcjk_destroy(h_); }
    GpuHF_IBS(const GpuHF_IBS&) = delete;
    GpuHF_IBS& operator=(const GpuHF_IBS&) = delete;

    qcjk_handle* handle() const { return h_; }
    int nbf() const { return n_; }
    const qcjk_stats& last_stats() const { return stats_; }
    // Shell quartets computed over every build so far, and those evaluated in FP32.
    long long quartets_total() const { return quartets_; }
    long long quartets_fp32_total() const { return quartets_fp32_; }
    int builds() const { return nbuilds_; }
    void set_flags(unsigned f) { flags_ = f; memo_valid_ = false; }

    // ---- WholeSystemFock_IBS ----
    size_t AODimension() const override { return (size_t)n_; }
    void AddAODensity(Mat& Dao, const Mat& D) const override { Dao += D; }  // identity O (one irrep)
    void SliceAOFock(Mat& Fab, const Mat& Fao) const override { Fab += Fao; }
    Mat MakeAOFock(const Mat& Dao, bool exchange) const override
    {
        ensure(Dao);
        return exchange ? K_ : J_;
    }

    // ---- Orbital_HF_IBS ----
    void AccumulateDirect(Mat& Sab, const Mat& Dcd, const Orbital_HF_IBS<double>* bs_cd) const override
    {
        check_partner(bs_cd);
        ensure(Dcd);
        Sab += J_;
    }
    void AccumulateExchange(Mat& Sab, const Mat& Dcd, const Orbital_HF_IBS<double>* bs_cd) const override
    {
        check_partner(bs_cd);
        ensure(Dcd);
        Sab += K_;
    }
    void AccumulateDirectBoth(Mat&, Mat&, const Mat&, const Mat&, const Orbital_HF_IBS<double>*) const override
    {
        // An irrep PAIR cannot arise: the Cartesian molecular basis is one irrep.
        throw std::runtime_error("GpuHF_IBS: the molecular Cartesian basis has one irrep; no irrep pair exists");
    }
    void AccumulateExchangeBoth(Mat&, Mat&, const Mat&, const Mat&, const Orbital_HF_IBS<double>*) const override
    {
        throw std::runtime_error("GpuHF_IBS: the molecular Cartesian basis has one irrep; no irrep pair exists");
    }

private:
    void check_partner(const Orbital_HF_IBS<double>* cd) const
    {
        if (cd != this)
            throw std::runtime_error("GpuHF_IBS: the HF partner basis is not this engine's basis");
    }
    void ensure(const Mat& D) const
    {
        if (D.n != n_) throw std::runtime_error("GpuHF_IBS: density has the wrong dimension");
        if (memo_valid_ && D.a.size() == memoD_.size() &&
            std::memcmp(D.a.data(), memoD_.data(), D.a.size() * sizeof(double)) == 0)
            return;
        J_ = Mat(n_);
        K_ = Mat(n_);
        const int rc = qcjk_build(h_, D.a.data(), (flags_ & QCJK_J) ? J_.a.data() : nullptr,
                                  (flags_ & QCJK_K) ? K_.a.data() : nullptr, flags_, &stats_);
        if (rc != QCJK_OK) throw EngineError("qcjk_build", rc);
        memoD_ = D.a;
        memo_valid_ = true;
        ++nbuilds_;
        quartets_ += stats_.quartets_computed;
        quartets_fp32_ += stats_.quartets_fp32;
    }

    qcjk_handle* h_ = nullptr;
    int n_ = 0;
    unsigned flags_;
    mutable bool memo_valid_ = false;
    mutable std::vector<double> memoD_;
    mutable Mat J_, K_;
    mutable qcjk_stats stats_{};
    mutable int nbuilds_ = 0;
    mutable long long quartets_ = 0, quartets_fp32_ = 0;
};

}  // namespace qchem_amd
