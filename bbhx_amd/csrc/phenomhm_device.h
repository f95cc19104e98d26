// PhenomD / PhenomHM amplitude, phase and t(f) on gfx950.
//
// The work is split in two so that the per-(binary, mode) coefficient setup
// (about 3k FP64 flops, ~40 transcendentals) is done once by one lane instead
// of once per frequency:
//   phm_setup(...)  fills a PhmRec for one (binary, mode)      [one lane]
//   phm_eval(...)   evaluates amp, phase, t(f) at one frequency [one lane per point]
//
// Physics and calibrated coefficients follow the reference's
// cutils/PhenomHMWaveform.cu (LALSimIMRPhenomD/HM; arXiv:1508.07250,
// 1508.07253, 1708.00404, QNM fits arXiv:1810.03550); the numbers of the fits
// are physics tables and are necessarily identical.
#pragma once

#include "common.h"
#include "dual.h"

namespace bbhx {
namespace phm {

// Non-deduced parameter type: the scalar type of a template below comes from
// its record argument; plain doubles passed elsewhere convert to it.
template <class T>
struct nd_t {
    using type = T;
};
template <class T>
using nd = typename nd_t<T>::type;

constexpr double AMP_JOIN_INS = 0.014;  // PhenomHMWaveform.hh:62
constexpr double PHI_JOIN_INS = 0.018;  // PhenomHMWaveform.hh:68

// ---------------------------------------------------------------- fits
// Calibrated fits a0 + a1 eta + (b0 + b1 eta + b2 eta^2) xi + (c..) xi^2 + (d..) xi^3
// with xi = chi_PN - 1 (arXiv:1508.07253 Table 5; PhenomHMWaveform.cu:390-1132).
enum FitId {
    F_RHO1, F_RHO2, F_RHO3, F_GAMMA1, F_GAMMA2, F_GAMMA3, F_V2,
    F_ALPHA1, F_ALPHA2, F_ALPHA3, F_ALPHA4, F_ALPHA5, F_BETA1, F_BETA2, F_BETA3,
    F_SIGMA1, F_SIGMA2, F_SIGMA3, F_SIGMA4, F_COUNT
};

__device__ const double FITS[F_COUNT][11] = {
    {3931.8979897196696, -17395.758706812805, 3132.375545898835, 343965.86092361377, -1.2162565819981997e6,
     -70698.00600428853, 1.383907177859705e6, -3.9662761890979446e6, -60017.52423652596, 803515.1181825735,
     -2.091710365941658e6},
    {-40105.47653771657, 112253.0169706701, 23561.696065836168, -3.476180699403351e6, 1.137593670849482e7,
     754313.1127166454, -1.308476044625268e7, 3.6444584853928134e7, 596226.612472288, -7.4277901143564405e6,
     1.8928977514040343e7},
    {83208.35471266537, -191237.7264145924, -210916.2454782992, 8.71797508352568e6, -2.6914942420669552e7,
     -1.9889806527362722e6, 3.0888029960154563e7, -8.390870279256162e7, -1.4535031953446497e6,
     1.7063528990822166e7, -4.2748659731120914e7},
    {0.006927402739328343, 0.03020474290328911, 0.006308024337706171, -0.12074130661131138, 0.26271598905781324,
     0.0034151773647198794, -0.10779338611188374, 0.27098966966891747, 0.0007374185938559283,
     -0.02749621038376281, 0.0733150789135702},
    {1.010344404799477, 0.0008993122007234548, 0.283949116804459, -4.049752962958005, 13.207828172665366,
     0.10396278486805426, -7.025059158961947, 24.784892370130475, 0.03093202475605892, -2.6924023896851663,
     9.609374464684983},
    {1.3081615607036106, -0.005537729694807678, -0.06782917938621007, -0.6689834970767117, 3.403147966134083,
     -0.05296577374411866, -0.9923793203111362, 4.820681208409587, -0.006134139870393713, -0.38429253308696365,
     1.7561754421985984},
    {0.8149838730507785, 2.5747553517454658, 1.1610198035496786, -2.3627771785551537, 6.771038707057573,
     0.7570782938606834, -2.7256896890432474, 7.1140380397149965, 0.1766934149293479, -0.7978690983168183,
     2.1162391502005153},
    {43.31514709695348, 638.6332679188081, -32.85768747216059, 2415.8938269370315, -5766.875169379177,
     -61.85459307173841, 2953.967762459948, -8986.29057591497, -21.571435779762044, 981.2158224673428,
     -3239.5664895930286},
    {-0.07020209449091723, -0.16269798450687084, -0.1872514685185499, 1.138313650449945, -2.8334196304430046,
     -0.17137955686840617, 1.7197549338119527, -4.539717148261272, -0.049983437357548705, 0.6062072055948309,
     -1.682769616644546},
    {9.5988072383479, -397.05438595557433, 16.202126189517813, -1574.8286986717037, 3600.3410843831093,
     27.092429659075467, -1786.482357315139, 5152.919378666511, 11.175710130033895, -577.7999423177481,
     1808.730762932043},
    {-0.02989487384493607, 1.4022106448583738, -0.07356049468633846, 0.8337006542278661, 0.2240008282397391,
     -0.055202870001177226, 0.5667186343606578, 0.7186931973380503, -0.015507437354325743, 0.15750322779277187,
     0.21076815715176228},
    {0.9974408278363099, -0.007884449714907203, -0.059046901195591035, 1.3958712396764088, -4.516631601676276,
     -0.05585343136869692, 1.7516580039343603, -5.990208965347804, -0.017945336522161195, 0.5965097794825992,
     -2.0608879367971804},
    {97.89747327985583, -42.659730877489224, 153.48421037904913, -1417.0620760768954, 2752.8614143665027,
     138.7406469558649, -1433.6585075135881, 2857.7418952430758, 41.025109467376126, -423.680737974639,
     850.3594335657173},
    {-3.282701958759534, -9.051384468245866, -12.415449742258042, 55.4716447709787, -106.05109938966335,
     -11.953044553690658, 76.80704618365418, -155.33172948098394, -3.4129261592393263, 25.572377569952536,
     -54.408036707740465},
    {-0.000025156429818799565, 0.000019750256942201327, -0.000018370671469295915, 0.000021886317041311973,
     0.00008250240316860033, 7.157371250566708e-6, -0.000055780000112270685, 0.00019142082884072178,
     5.447166261464217e-6, -0.00003220610095021982, 0.00007974016714984341},
    {2096.551999295543, 1463.7493168261553, 1312.5493286098522, 18307.330017082117, -43534.1440746107,
     -833.2889543511114, 32047.31997183187, -108609.45037520859, 452.25136398112204, 8353.439546391714,
     -44531.3250037322},
    {-10114.056472621156, -44631.01109458185, -6541.308761668722, -266959.23419307504, 686328.3229317984,
     3405.6372187679685, -437507.7208209015, 1.6318171307344697e6, -7462.648563007646, -114585.25177153319,
     674402.4689098676},
    {22933.658273436497, 230960.00814979506, 14961.083974183695, 1.1940181342318142e6, -3.1042239693052764e6,
     -3038.166617199259, 1.8720322849093592e6, -7.309145012085539e6, 42738.22871475411, 467502.018616601,
     -3.064853498512499e6},
    {-14621.71522218357, -377812.8579387104, -9608.682631509726, -1.7108925257214056e6, 4.332924601416521e6,
     -22366.683262266528, -2.5019716386377467e6, 1.0274495902259542e7, -85360.30079034246, -570025.3441737515,
     4.396844346849777e6},
};

template <class S>
__device__ inline S fit(int id, S eta, S chi)
{
    const double* k = FITS[id];
    S xi = -1 + chi;
    S xi2 = xi * xi, xi3 = xi2 * xi, eta2 = eta * eta;
    return k[0] + k[1] * eta + (k[2] + k[3] * eta + k[4] * eta2) * xi + (k[5] + k[6] * eta + k[7] * eta2) * xi2 +
           (k[8] + k[9] * eta + k[10] * eta2) * xi3;
}

// Powers of x built from x^(1/6) (PhenomHMWaveform.cu:354-368).
template <class S>
struct P6T {
    S sixth, third, two_thirds, four_thirds, five_thirds, two, seven_thirds, eight_thirds;
};

template <class S>
__device__ inline P6T<S> pow6(S x)
{
    P6T<S> p;
    p.sixth = root6_pos(x);
    p.third = p.sixth * p.sixth;
    p.two_thirds = p.third * p.third;
    p.four_thirds = x * p.third;
    p.five_thirds = p.four_thirds * p.third;
    p.two = x * x;
    p.seven_thirds = p.third * p.two;
    p.eight_thirds = p.two_thirds * p.two;
    return p;
}

// ------------------------------------------------------------- record

// Everything point evaluation needs for one (binary, mode). ~100 doubles.
template <class S>
struct PhmRecT {
    // binary level
    S M_sec, amp_scale, t0, phi0, Mf_ref;
    // leading-order PN amplitude of this mode (IMRPhenomHMOnePointFiveSpinPN):
    // |H(v)| = |(c1 v + c2 v^2 + c3 v^3 + c4 v^4) + i c4i v^4|
    S pn_c1, pn_c2, pn_c3, pn_c4, pn_c4i;
    S pn_vscale, pn_v2ratio;  // 2 pi M / m and (2/m)^(1/3) of this mode
    // PhenomD amplitude (mode independent)
    S a_fRD, a_fDM, a_g1, a_g2, a_g3, a_fmax;
    S a_rho1, a_rho2, a_rho3, a_d0, a_d1, a_d2, a_d3, a_d4;
    S a_pf23, a_pf1, a_pf43, a_pf53, a_pf2, a_amp0;
    // PhenomD phase with this mode's ringdown ratios
    S p_eta, p_etaInv, p_fRD, p_fDM, p_Rholm, p_Taulm;
    S beta1, beta2, beta3, alpha1, alpha2, alpha3, alpha4, alpha5;
    S sigma1, sigma2, sigma3, sigma4;
    S C1Int, C2Int, C1MRD, C2MRD, fMRDJoin;
    S pn_v0, pn_v2, pn_v3, pn_v4, pn_v5, pn_v6, pn_v7, pn_vl5, pn_vl6;
    S q_init, q_23, q_13, q_13log, q_logv, q_m13, q_m23, q_m1, q_m53, q_1, q_43, q_53, q_2;
    // HM frequency maps: phase (a*, b*, fi, fr) and amplitude (amp_*)
    S ai, bi, am, bm, ar, br, fi, fr, PhDBconst, PhDCconst, PhDBAterm, phaseC;
    S iai, iam, iar;  // 1 / ai, 1 / am, 1 / ar
    S amp_fi, amp_fr, amp_Ai, amp_Am, amp_Bm, amp_Br;
    S cshift;
    int ell, mm, phenomd, pad;
};
using PhmRec = PhmRecT<double>;

// ------------------------------------------------------------ amplitude

template <class S>
__device__ inline S amp_ins(const PhmRecT<S>& r, nd<S> Mf, const P6T<S>& p)
{
    S Mf2 = p.two, Mf3 = Mf * Mf2;
    return 1 + p.two_thirds * r.a_pf23 + Mf * r.a_pf1 + p.four_thirds * r.a_pf43 + p.five_thirds * r.a_pf53 +
           Mf2 * r.a_pf2 + p.seven_thirds * r.a_rho1 + p.eight_thirds * r.a_rho2 + Mf3 * r.a_rho3;
}

template <class S>
__device__ inline S amp_mrd(const PhmRecT<S>& r, nd<S> f)
{
    const S fDMg3 = r.a_fDM * r.a_g3;
    const S d = f - r.a_fRD;
    const S inv = rcp_pos(fDMg3);
    return exp_fast(-(d * r.a_g2) * inv) * (fDMg3 * r.a_g1) * rcp_pos(d * d + fDMg3 * fDMg3);
}

template <class S>
__device__ inline S amp_mrd_deriv(const PhmRecT<S>& r, nd<S> f)
{
    S fDMg3 = r.a_fDM * r.a_g3;
    S d = f - r.a_fRD;
    S e = exp((d * r.a_g2) / (fDMg3));
    S s = d * d + fDMg3 * fDMg3;
    return (-2 * r.a_fDM * (d)*r.a_g3 * r.a_g1) / (e * (s * s)) - (r.a_g2 * r.a_g1) / (e * (s));
}

// IMRPhenDAmplitude without the distance factor (PhenomHMWaveform.cu:1508-1537)
template <class S>
__device__ inline S amp_imr(const PhmRecT<S>& r, nd<S> f, const P6T<S>& p)
{
    const S pre = r.a_amp0 * rcp_pos(f * p.sixth);
    if (!(f >= AMP_JOIN_INS)) return pre * amp_ins(r, f, p);
    if (f >= r.a_fmax) return pre * amp_mrd(r, f);
    S f2 = f * f, f3 = f * f2, f4 = f * f3;
    return pre * (r.a_d0 + r.a_d1 * f + r.a_d2 * f2 + r.a_d3 * f3 + r.a_d4 * f4);
}

// The coefficients of |H_lm(v)|, the leading-order PN amplitude
// (IMRPhenomHMOnePointFiveSpinPN, PhenomHMWaveform.cu:2405-2449) without the
// common v^-3.5 prefactor, as a polynomial in v per (binary, mode): c[0..3] of
// v..v^4 in the real part, c[4] of v^4 in the imaginary part. Modes other
// than (2,1), (3,3), (3,2), (4,4), (4,3) have none (their factor is 1).
template <class S>
__device__ inline void pn_habs_coeffs(int l, int m, S eta, S delta, S Xs, S Xa, S c[5])
{
    for (int k = 0; k < 5; k++) c[k] = 0.0;
    if (l == 2 && m == 1) {
        const double pre = 0.47140452079103168;  // sqrt(2) / 3
        c[0] = pre * delta;
        c[1] = pre * (-1.5 * (Xa + delta * Xs));
        c[2] = pre * (delta * ((335.0 / 672.0) + (eta * 117.0 / 56.0)));
        c[3] = pre * (Xa * (3427.0 / 1344 - eta * 2101.0 / 336) + delta * Xs * (3427.0 / 1344 - eta * 965 / 336) +
                      delta * (-PI));
        c[4] = pre * (delta * (-0.5 - 2.0 * 0.69314718056));
    } else if (l == 3 && m == 3) {
        c[0] = 0.75 * sqrt(5.0 / 7.0) * delta;
    } else if (l == 3 && m == 2) {
        c[1] = (1.0 / 3.0) * sqrt(5.0 / 7.0) * (1.0 - 3.0 * eta);
    } else if (l == 4 && m == 4) {
        c[1] = (4.0 / 9.0) * sqrt(10.0 / 7.0) * (1.0 - 3.0 * eta);
    } else if (l == 4 && m == 3) {
        c[2] = 0.75 * sqrt(3.0 / 35.0) * delta * (1.0 - 2.0 * eta);
    }
}

// |H_lm(v)| from the record's coefficients: one polynomial for every mode, so
// the evaluation has no branch on the mode
template <class S>
__device__ inline S pn_habs2(const PhmRecT<S>& r, nd<S> v)
{
    const S re = v * (r.pn_c1 + v * (r.pn_c2 + v * (r.pn_c3 + v * r.pn_c4)));
    const S v2 = v * v;
    const S im = (v2 * v2) * r.pn_c4i;
    return re * re + im * im;
}

// The HM amplitude factor beta * HM1 / HM2 of get_amp (PhenomHMWaveform.cu:2473-2511).
// With v1 = (2 pi M Mf/m)^(1/3), v2 = (2/m)^(1/3) v1, v3 = (2 pi M Mf_amp/m)^(1/3) and
// v4 = (m/2)^(1/3) v3, the v^-3.5 prefactors of the four PN amplitudes cancel
// exactly, leaving |H(v1)| |H(v3)| / |H(v2)| (1 for the (2,2) mode; 0 where the
// reference's beta_term1 vanishes).
//
// Below amp_fi the amplitude map is Mf_amp = (2/m) Mf (PhenomHMWaveform.cu:2200-2212),
// so v3 = (2/m)^(1/3) v1 = v2 and the factor is |H(v1)| exactly.
template <class S>
__device__ inline S hm_amp_factor(const PhmRecT<S>& r, nd<S> Mf_wf, nd<S> Mf_amp)
{
    if (r.ell == 2 && r.mm == 2) return 1.0;
    const S v1 = cbrt_pos(r.pn_vscale * Mf_wf);
    const S h1 = pn_habs2(r, v1);  // |H(v1)|^2
    if (h1 == 0.0) return 0.0;
    if (!(Mf_wf > r.amp_fi)) return sqrt(h1);
    const S v2 = v1 * r.pn_v2ratio;
    const S v3 = cbrt_pos(r.pn_vscale * Mf_amp);
    // |H(v2)| > 0 wherever |H(v1)| > 0 (the same factors of eta and delta
    // vanish); one square root of the squared magnitudes' ratio
    return sqrt(h1 * pn_habs2(r, v3) * rcp_pos(pn_habs2(r, v2)));
}

// ---------------------------------------------------------------- phase

// Point evaluations take iMf = 1/Mf (computed once per point): every negative
// power below is a product of iMf and the positive powers, so the only
// per-point divisions are that reciprocal and 1/v.

template <class S>
__device__ inline S phi_ins(const PhmRecT<S>& r, nd<S> Mf, const P6T<S>& p, nd<S> logv, nd<S> iMf)
{
    const S im13 = p.two_thirds * iMf;  // Mf^(-1/3)
    const S im23 = im13 * im13;
    S ph = r.q_init;
    ph += r.q_23 * p.two_thirds;
    ph += r.q_13 * p.third;
    ph += r.q_13log * logv * p.third;
    ph += r.q_logv * logv;
    ph += r.q_m13 * im13;
    ph += r.q_m23 * im23;
    ph += r.q_m1 * iMf;
    ph += r.q_m53 * (im23 * iMf);
    ph += (r.q_1 * Mf + r.q_43 * p.four_thirds + r.q_53 * p.five_thirds + r.q_2 * p.two) * r.p_etaInv;
    return ph;
}

// d phi_ins / dMf with v = (pi Mf)^(1/3) given
template <class S>
__device__ inline S dphi_ins(const PhmRecT<S>& r, nd<S> v, nd<S> logv)
{
    const S v2 = v * v, v3 = v * v2, v4 = v * v3, v5 = v * v4, v6 = v * v5, v7 = v * v6;
    const S iv = rcp_pos(v), iv2 = iv * iv, iv4 = iv2 * iv2, iv8 = iv4 * iv4;
    S D = 0.0;
    D += +2.0 * r.pn_v7 * v7;
    D += (r.pn_v6 + r.pn_vl6 * (1.0 + logv)) * v6;
    D += r.pn_vl5 * v5;
    D += -1.0 * r.pn_v4 * v4;
    D += -2.0 * r.pn_v3 * v3;
    D += -3.0 * r.pn_v2 * v2;
    D += -5.0 * r.pn_v0;
    D *= (PI / 3.0) * iv8;
    // pi^(-1/3), pi^(-2/3), 1/pi
    const double ipi13 = 0.68278406325529568, ipi23 = 0.46619407703541166, ipi = 0.31830988618379067;
    D += (r.sigma1 + r.sigma2 * v * ipi13 + r.sigma3 * v2 * ipi23 + (r.sigma4 * ipi) * v3) * r.p_etaInv;
    return D;
}

template <class S>
__device__ inline S phi_int(const PhmRecT<S>& r, nd<S> Mf, nd<S> iMf, nd<S> logMf)
{
    return r.beta1 * Mf - r.beta3 * (iMf * iMf * iMf) * (1.0 / 3.0) + r.beta2 * logMf;
}

template <class S>
__device__ inline S dphi_int(const PhmRecT<S>& r, nd<S> iMf)
{
    const S im2 = iMf * iMf;
    return (r.beta1 + r.beta3 * (im2 * im2) + r.beta2 * iMf) * r.p_etaInv;
}

template <class S>
__device__ inline S phi_mrd(const PhmRecT<S>& r, nd<S> f, nd<S> iMf)
{
    const S sf = sqrt(f);
    const S f075 = sqrt(f * sf);
    return -(r.alpha2 * iMf) + (4.0 / 3.0) * (r.alpha3 * f075) + r.alpha1 * f +
           r.alpha4 * r.p_Rholm * atan_fast((f - r.alpha5 * r.p_fRD) * rcp_pos(r.p_Rholm * r.p_fDM * r.p_Taulm));
}

template <class S>
__device__ inline S dphi_mrd(const PhmRecT<S>& r, nd<S> f, nd<S> iMf)
{
    const S num = f - r.alpha5 * r.p_fRD;
    const S q = num * rcp_pos(r.p_fDM * r.p_Taulm * r.p_Rholm);
    const S f075 = sqrt(f * sqrt(f));  // f^(-1/4) = f^(3/4) / f
    return (r.alpha1 + r.alpha2 * (iMf * iMf) + r.alpha3 * (f075 * iMf) +
            r.alpha4 * rcp_pos(r.p_fDM * r.p_Taulm * (1 + q * q))) *
           r.p_etaInv;
}

// IMRPhenDPhase and its derivative at the same Mf, sharing the powers. With
// 'same' set (the lane's f equals the one 'pre' was built from), the inspiral
// branch reuses the powers 'pre' instead of computing them again. One log(Mf)
// serves both the inspiral (log v = (log Mf + log pi) / 3) and the intermediate
// piece, so a wavefront whose lanes straddle their join computes it once.
template <class S>
__device__ inline void phase_imr(const PhmRecT<S>& r, nd<S> f, S& ph, S* dph, const P6T<S>* pre = nullptr,
                                 bool same = false)
{
    const S iMf = rcp_pos(f);
    const bool ins = !(f >= PHI_JOIN_INS);
    if (!ins && f >= r.fMRDJoin) {
        ph = r.p_etaInv * phi_mrd(r, f, iMf) + r.C1MRD + r.C2MRD * f;
        if (dph) *dph = dphi_mrd(r, f, iMf) + r.C2MRD;
        return;
    }
    const S logMf = log_pos(f);
    if (ins) {
        P6T<S> p;
        if (pre && same) p = *pre;
        else p = pow6(f);
        const double pi13 = 1.4645918875615231, LOG_PI = 1.1447298858494002;
        S v = p.third * pi13;
        S logv = (logMf + LOG_PI) * (1.0 / 3.0);
        ph = phi_ins(r, f, p, logv, iMf);
        if (dph) *dph = dphi_ins(r, v, logv);
        return;
    }
    ph = r.p_etaInv * phi_int(r, f, iMf, logMf) + r.C1Int + r.C2Int * f;
    if (dph) *dph = dphi_int(r, iMf) + r.C2Int;
}

template <class S>
__device__ inline S phase_only(const PhmRecT<S>& r, nd<S> f)
{
    S ph;
    phase_imr(r, f, ph, (S*)nullptr);
    return ph;
}

// -------------------------------------------------------- final state / QNM

template <class S>
__device__ inline void enforce_primary(S& m1, S& m2, S& c1, S& c2)
{
    if (!(m1 > m2)) {
        S t = m1; m1 = m2; m2 = t;
        t = c1; c1 = c2; c2 = t;
    }
}

// IMRPhenomDFinalMass / FinalSpin (PhenomHMWaveform.cu:1679-1814)
template <class S>
__device__ inline void final_state(S m1, S m2, S chi1, S chi2, S& finmass, S& finspin)
{
    enforce_primary(m1, m2, chi1, chi2);
    S M = m1 + m2;
    S eta = m1 * m2 / (M * M);
    S Seta = sqrt(1.0 - 4.0 * eta);
    S q1 = 0.5 * (1.0 + Seta), q2 = 0.5 * (1.0 - Seta);
    S q1s = q1 * q1, q2s = q2 * q2;
    S eta2 = eta * eta, eta3 = eta2 * eta, eta4 = eta3 * eta;
    S s = (q1s * chi1 + q2s * chi2) / (q1s + q2s);
    S erad = ((0.055974469826360077 * eta + 0.5809510763115132 * eta2 - 0.9606726679372312 * eta3 +
                    3.352411249771192 * eta4) *
                   (1. + (-0.0030302335878845507 - 2.0066110851351073 * eta + 7.7050567802399215 * eta2) * s)) /
                  (1. + (-0.6714403054720589 - 1.4756929437702908 * eta + 7.304676214885011 * eta2) * s);
    finmass = 1.0 - erad;
    S sf = (q1s * chi1 + q2s * chi2);
    S sf2 = sf * sf, sf3 = sf2 * sf, sf4 = sf3 * sf;
    finspin = 3.4641016151377544 * eta - 4.399247300629289 * eta2 + 9.397292189321194 * eta3 -
              13.180949901606242 * eta4 + (1 - 0.0850917821418767 * eta - 5.837029316602263 * eta2) * sf +
              (0.1014665242971878 * eta - 2.0967746996832157 * eta2) * sf2 +
              (-1.3546806617824356 * eta + 4.108962025369336 * eta2) * sf3 +
              (-0.8676969352555539 * eta + 2.064046835273906 * eta2) * sf4;
}

// Complex QNM frequency fit CW07102016 (n = 0) of the (l, m) mode in units of 1/M_final;
// returns false for a mode without a fit. (PhenomHMWaveform.cu:56-211)
template <class S>
__device__ inline bool qnm_cw(S kappa, int l, int m_in, S& out_re, S& out_im)
{
    int m = m_in < 0 ? -m_in : m_in;
    // complex coefficient per power of kappa, A_i e^{i P_i} of the fits in
    // Cartesian form (A_i cos P_i, A_i sin P_i; the published amplitude / phase
    // pairs, evaluated once so that no sine or cosine is taken per record);
    // 'base' is the kappa^0 real offset for zero-damped modes, whose series
    // starts at kappa^1
    const double* C;
    const double* Sn;
    int n;
    double base;
    bool zd;
    static constexpr double C22[5] = {-1.5137612398456926, 1.82438009537132, -1.9491472924851778, 1.3160523687197017,
                                      -0.386253596520847};
    static constexpr double S22[5] = {0.36798666300537497, -0.6913198959256075, 0.7807925359505982,
                                      -0.5087007674840115, 0.13940135002826776};
    static constexpr double C32[7] = {1.0224518751857428, 0.19456949695384276, -1.7046730536276729, 0.9334708870357518,
                                      1.2817015416050352, -2.047141684957698, 0.8331491642591967};
    static constexpr double S32[7] = {0.0049793799973692964, 0.15266302925156916, 0.00564189789146444,
                                      0.15375458653347265, -0.8390203896453422, 1.0049939601312927,
                                      -0.3937391358806763};
    static constexpr double C44[5] = {-2.6333344971899724, 2.8983059270065543, -3.1068517011428427, 2.036655322665644,
                                      -0.5743326515202851};
    static constexpr double S44[5] = {0.3678874534055692, -0.6854411640438772, 0.8400728137834692, -0.6155714539424388,
                                      0.18496398129024574};
    static constexpr double C21[7] = {0.5885550727060468, -0.12446771988120166, 1.0237846421677674, -5.5694819533803495,
                                      10.070990838680922, -8.324170006013349, 2.678477439380642};
    static constexpr double S21[7] = {0.02563304821865536, 0.14217947242002135, -0.5240832169327109, 2.352281987769593,
                                      -4.729608920661527, 4.251675580693241, -1.434409619054415};
    static constexpr double C33[5] = {-2.06305539331511, 2.3690119585178913, -2.526835645169051, 1.6514625332798913,
                                      -0.46707723604718027};
    static constexpr double S33[5] = {0.36821285387997693, -0.6977975754308797, 0.8486034802921025, -0.6037524077483988,
                                      0.17577435831730695};
    static constexpr double C43[5] = {0.1697708308937892, -3.07896855783127, 4.110293599941198, -2.8795369847373125,
                                      0.8476113318527659};
    static constexpr double S43[5] = {0.11498576909179864, 0.3881163320136248, -1.0247991220739723, 0.9394221404554335,
                                      -0.32682921703275014};
    if (l == 2 && m == 2) {
        C = C22; Sn = S22; n = 5; base = 1.0; zd = true;
    } else if (l == 3 && m == 2) {
        C = C32; Sn = S32; n = 7; base = 0.0; zd = false;
    } else if (l == 4 && m == 4) {
        C = C44; Sn = S44; n = 5; base = 2.0; zd = true;
    } else if (l == 2 && m == 1) {
        C = C21; Sn = S21; n = 7; base = 0.0; zd = false;
    } else if (l == 3 && m == 3) {
        C = C33; Sn = S33; n = 5; base = 1.5; zd = true;
    } else if (l == 4 && m == 3) {
        C = C43; Sn = S43; n = 5; base = 1.5; zd = true;
    } else {
        out_re = 0.0;
        out_im = 0.0;
        return false;
    }
    S sum_re = 0.0, sum_im = 0.0;
    S kp = 1.0;
    for (int i = 0; i < n; i++) {
        sum_re += C[i] * kp;
        sum_im += Sn[i] * kp;
        kp *= kappa;
    }
    S ans_re = zd ? base + kappa * sum_re : sum_re;
    S ans_im = zd ? kappa * sum_im : sum_im;
    if (m_in < 0) ans_re = -ans_re;
    out_re = ans_re;
    out_im = ans_im;
    return true;
}

template <class S>
__device__ inline bool hm_ringdown(int ell, int mm, S finmass, S finspin, S& frd, S& fdm)
{
    int am = mm < 0 ? -mm : mm;
    S kappa = pow(log(2.0 - finspin) / log(3.0), 1.0 / (2.0 + ell - am));
    S zr, zi;
    bool ok = qnm_cw(kappa, ell, mm, zr, zi);
    const double inv2pi = 0.5 / PI;
    frd = inv2pi * zr / finmass;
    fdm = inv2pi * zi / finmass;
    return ok;
}

// ---------------------------------------------------------------- setup

// TaylorF2 + PhenomD phase coefficients + C1 joins; amplitude coefficients when
// 'with_amp' (PhenomHMWaveform.cu:253-335, :446-874, :1261-1339, :1587-1644).
// The primary is enforced on the copies again, as the reference does.
template <class S>
__device__ inline void setup_phenomd(PhmRecT<S>& r, nd<S> m1, nd<S> m2, nd<S> chi1, nd<S> chi2, nd<S> Rholm,
                                     nd<S> Taulm, nd<S> fring, nd<S> fdamp, bool with_amp)
{
    enforce_primary(m1, m2, chi1, chi2);
    const S Mtot = m1 + m2;
    const S eta = m1 * m2 / (Mtot * Mtot);
    const S Seta = sqrt(1.0 - 4.0 * eta);
    const S chi = (chi1 + chi2) / 2.0 * (1.0 - eta * 76.0 / 113.0) + Seta * ((chi1 - chi2) / 2.0);

    r.p_eta = eta;
    r.p_etaInv = 1.0 / eta;
    r.p_fRD = fring;
    r.p_fDM = fdamp;
    r.p_Rholm = Rholm;
    r.p_Taulm = Taulm;
    r.sigma1 = fit(F_SIGMA1, eta, chi);
    r.sigma2 = fit(F_SIGMA2, eta, chi);
    r.sigma3 = fit(F_SIGMA3, eta, chi);
    r.sigma4 = fit(F_SIGMA4, eta, chi);
    r.beta1 = fit(F_BETA1, eta, chi);
    r.beta2 = fit(F_BETA2, eta, chi);
    r.beta3 = fit(F_BETA3, eta, chi);
    r.alpha1 = fit(F_ALPHA1, eta, chi);
    r.alpha2 = fit(F_ALPHA2, eta, chi);
    r.alpha3 = fit(F_ALPHA3, eta, chi);
    r.alpha4 = fit(F_ALPHA4, eta, chi);
    r.alpha5 = fit(F_ALPHA5, eta, chi);

    // TaylorF2 aligned-spin phasing minus the 3PN spin-spin term
    {
        const S d = (m1 - m2) / (m1 + m2);
        const S eta_pn = m1 * m2 / Mtot / Mtot;
        const S m1M = m1 / Mtot, m2M = m2 / Mtot;
        const S SL = m1M * m1M * chi1 + m2M * m2M * chi2;
        const S dSigmaL = d * (m2M * chi2 - m1M * chi1);
        const S pfaN = 3. / (128. * eta_pn);
        const S e = eta_pn;
        S v0 = 1. * pfaN;
        S v2 = pfaN * (5. * (743. / 84. + 11. * e) / 9.);
        S v3 = pfaN * (-16. * PI);
        S v4 = pfaN * (5. * (3058.673 / 7.056 + 5429. / 7. * e + 617. * e * e) / 72.);
        S v5 = pfaN * (5. / 9. * (7729. / 84. - 13. * e) * PI);
        S vl5 = pfaN * (5. / 3. * (7729. / 84. - 13. * e) * PI);
        S v6 = pfaN * ((11583.231236531 / 4.694215680 - 640. / 3. * PI * PI - 6848. / 21. * GAMMA_E) +
                            e * (-15737.765635 / 3.048192 + 2255. / 12. * PI * PI) + e * e * 76055. / 1728. -
                            e * e * e * 127825. / 1296.);
        v6 += pfaN * (-6848. / 21.) * 1.3862943611198906;  // log(4)
        S vl6 = pfaN * (-6848. / 21.);
        S v7 = pfaN * (PI * (77096675. / 254016. + 378515. / 1512. * e - 74045. / 756. * e * e));
        S sig = e * (721. / 48. * chi1 * chi2 - 247. / 48. * (chi1 * chi2));
        sig += (720. - 1.) / 96.0 * m1M * m1M * chi1 * chi1;
        sig += (720. - 1.) / 96.0 * m2M * m2M * chi2 * chi2;
        sig -= (240. - 7.) / 96.0 * m1M * m1M * (chi1 * chi1);
        sig -= (240. - 7.) / 96.0 * m2M * m2M * (chi2 * chi2);
        S ss1 = (4703.5 / 8.4 + 2935. / 6. * m1M - 120. * m1M * m1M) + (-4108.25 / 6.72 - 108.5 / 1.2 * m1M + 125.5 / 3.6 * m1M * m1M);
        S ss2 = (4703.5 / 8.4 + 2935. / 6. * m2M - 120. * m2M * m2M) + (-4108.25 / 6.72 - 108.5 / 1.2 * m2M + 125.5 / 3.6 * m2M * m2M);
        S ss3 = (326.75 / 1.12 + 557.5 / 1.8 * e) * e * chi1 * chi2;
        ss3 += ss1 * m1M * m1M * (chi1 * chi1);
        ss3 += ss2 * m2M * m2M * (chi2 * chi2);
        S gam = (554345. / 1134. + 110. * e / 9.) * SL + (13915. / 84. - 10. * e / 3.) * dSigmaL;
        v7 += pfaN * ((-8980424995. / 762048. + 6586595. * e / 756. - 305. * e * e / 36.) * SL -
                      (170978035. / 48384. - 2876425. * e / 672. - 4735. * e * e / 144.) * dSigmaL);
        v6 += pfaN * (PI * (3760. * SL + 1490. * dSigmaL) / 3. + ss3);
        v5 += pfaN * (-1. * gam);
        vl5 += pfaN * (-3. * gam);
        v4 += pfaN * (-10. * sig);
        v3 += pfaN * (188. * SL / 3. + 25. * dSigmaL);
        v6 -= ss3 * v0;
        r.pn_v0 = v0; r.pn_v2 = v2; r.pn_v3 = v3; r.pn_v4 = v4; r.pn_v5 = v5;
        r.pn_v6 = v6; r.pn_v7 = v7; r.pn_vl5 = vl5; r.pn_vl6 = vl6;
    }

    const P6T<double> pp = pow6(PI);
    r.q_init = r.pn_v5 - PI_4;
    r.q_23 = r.pn_v7 * pp.two_thirds;
    r.q_13 = r.pn_v6 * pp.third;
    r.q_13log = r.pn_vl6 * pp.third;
    r.q_logv = r.pn_vl5;
    r.q_m13 = r.pn_v4 / pp.third;
    r.q_m23 = r.pn_v3 / pp.two_thirds;
    r.q_m1 = r.pn_v2 / PI;
    r.q_m53 = r.pn_v0 / pp.five_thirds;
    r.q_1 = r.sigma1;
    r.q_43 = r.sigma2 * 3.0 / 4.0;
    r.q_53 = r.sigma3 * 3.0 / 5.0;
    r.q_2 = r.sigma4 / 2.0;

    // C^1 joins (PhenomHMWaveform.cu:1303-1339)
    r.fMRDJoin = 0.5 * r.p_fRD;
    r.C1Int = 0.0;
    r.C2Int = 0.0;
    {
        const S ij = 1.0 / PHI_JOIN_INS;
        S vj = cbrt(PI * PHI_JOIN_INS);
        S dins = dphi_ins(r, vj, log(vj));
        r.C2Int = dins - dphi_int(r, ij);
        P6T<S> pj = pow6(S(PHI_JOIN_INS));
        const double logj = log_pos(PHI_JOIN_INS);
        S logvj = (logj + 1.1447298858494002) * (1.0 / 3.0);
        r.C1Int = phi_ins(r, PHI_JOIN_INS, pj, logvj, ij) - r.p_etaInv * phi_int(r, PHI_JOIN_INS, ij, logj) -
                  r.C2Int * PHI_JOIN_INS;
        const S fj = r.fMRDJoin;
        const S ifj = 1.0 / fj;
        S phi_int_j = r.p_etaInv * phi_int(r, fj, ifj, log_pos(fj)) + r.C1Int + r.C2Int * fj;
        S fj2 = fj * fj;
        S dphi_int_j = r.C2Int + (r.beta1 + r.beta3 / (fj2 * fj2) + r.beta2 / fj) / r.p_eta;
        r.C2MRD = dphi_int_j - dphi_mrd(r, fj, ifj);
        r.C1MRD = phi_int_j - r.p_etaInv * phi_mrd(r, fj, ifj) - r.C2MRD * fj;
    }

    if (!with_amp) return;

    // amplitude coefficients (PhenomHMWaveform.cu:446-874)
    r.a_fRD = fring;
    r.a_fDM = fdamp;
    r.a_g1 = fit(F_GAMMA1, eta, chi);
    r.a_g2 = fit(F_GAMMA2, eta, chi);
    r.a_g3 = fit(F_GAMMA3, eta, chi);
    if (r.a_g2 <= 1)
        r.a_fmax = fabs(r.a_fRD + (r.a_fDM * (-1 + sqrt(1 - r.a_g2 * r.a_g2)) * r.a_g3) / r.a_g2);
    else
        r.a_fmax = fabs(r.a_fRD + (r.a_fDM * (-1) * r.a_g3) / r.a_g2);
    r.a_rho1 = fit(F_RHO1, eta, chi);
    r.a_rho2 = fit(F_RHO2, eta, chi);
    r.a_rho3 = fit(F_RHO3, eta, chi);

    const S c12 = chi1 * chi1, c22 = chi2 * chi2, eta2 = eta * eta, eta3 = eta * eta2;
    const S Pi = PI, Pi2 = pp.two;
    r.a_amp0 = (sqrt(2.0 / 3.0) * sqrt(eta)) / pow(PI, 1. / 6.);
    r.a_pf23 = ((-969 + 1804 * eta) * pp.two_thirds) / 672.;
    r.a_pf1 = ((chi1 * (81 * (1 + Seta) - 44 * eta) + chi2 * (81 - 81 * Seta - 44 * eta)) * Pi) / 48.;
    const S k43 = (-27312085.0 - 10287648 * c22 - 10287648 * c12 * (1 + Seta) + 10287648 * c22 * Seta +
                        24 * (-1975055 + 857304 * c12 - 994896 * chi1 * chi2 + 857304 * c22) * eta + 35371056 * eta2);
    r.a_pf43 = (k43 * pp.four_thirds) / 8.128512e6;
    const S k53 = (chi2 * (-285197 * (-1 + Seta) + 4 * (-91902 + 1579 * Seta) * eta - 35632 * eta2) +
                        chi1 * (285197 * (1 + Seta) - 4 * (91902 + 1579 * Seta) * eta - 35632 * eta2) +
                        42840 * (-1.0 + 4 * eta) * Pi);
    r.a_pf53 = (pp.five_thirds * k53) / 32256.;
    const S k2 = (-336 * (-3248849057.0 + 2943675504 * c12 - 3339284256 * chi1 * chi2 + 2943675504 * c22) * eta2 -
                       324322727232 * eta3 -
                       7 * (-177520268561 + 107414046432 * c22 + 107414046432 * c12 * (1 + Seta) - 107414046432 * c22 * Seta +
                            11087290368 * (chi1 + chi2 + chi1 * Seta - chi2 * Seta) * Pi) +
                       12 * eta * (-545384828789 - 176491177632 * chi1 * chi2 + 202603761360 * c22 +
                                   77616 * c12 * (2610335 + 995766 * Seta) - 77287373856 * c22 * Seta +
                                   5841690624 * (chi1 + chi2) * Pi + 21384760320 * Pi2));
    r.a_pf2 = -(Pi2 * k2) / 6.0085960704e10;

    // intermediate amplitude by collocation (PhenomHMWaveform.cu:656-842)
    const S f1 = AMP_JOIN_INS, f3 = r.a_fmax;
    const S f2 = f1 + (f3 - f1) / 2.0;
    const P6T<S> p1 = pow6(S(f1));
    const S v1 = amp_ins(r, f1, p1);
    // d/dMf of the inspiral ansatz at f1 (PhenomHMWaveform.cu:487-506)
    const S d1 = ((-969 + 1804 * eta) * pow(Pi, 2.0 / 3.0)) / (1008. * pow(f1, 1.0 / 3.0)) +
                      ((chi1 * (81 * (1 + Seta) - 44 * eta) + chi2 * (81 - 81 * Seta - 44 * eta)) * Pi) / 48. +
                      (k43 * pow(f1, 1.0 / 3.0) * pow(Pi, 4.0 / 3.0)) / 6.096384e6 +
                      (5 * pow(f1, 2.0 / 3.0) * pow(Pi, 5.0 / 3.0) * k53) / 96768. -
                      (f1 * (Pi * Pi) * k2) / 3.0042980352e10 + (7.0 / 3.0) * pow(f1, 4.0 / 3.0) * r.a_rho1 +
                      (8.0 / 3.0) * pow(f1, 5.0 / 3.0) * r.a_rho2 + 3 * (f1 * f1) * r.a_rho3;
    const S v3 = amp_mrd(r, f3);
    const S d2 = amp_mrd_deriv(r, f3);
    const S v2 = fit(F_V2, eta, chi);
    const S f12 = f1 * f1, f13 = f1 * f12, f14 = f1 * f13, f15 = f1 * f14;
    const S f22 = f2 * f2, f23 = f2 * f22, f24 = f2 * f23;
    const S f32 = f3 * f3, f33 = f3 * f32, f34 = f3 * f33, f35 = f3 * f34;
    const S den = ((f1 - f2) * (f1 - f2)) * ((f1 - f3) * (f1 - f3) * (f1 - f3)) * ((f3 - f2) * (f3 - f2));
    r.a_d0 = -((d2 * f15 * f22 * f3 - 2 * d2 * f14 * f23 * f3 + d2 * f13 * f24 * f3 - d2 * f15 * f2 * f32 + d2 * f14 * f22 * f32 -
                d1 * f13 * f23 * f32 + d2 * f13 * f23 * f32 + d1 * f12 * f24 * f32 - d2 * f12 * f24 * f32 + d2 * f14 * f2 * f33 +
                2 * d1 * f13 * f22 * f33 - 2 * d2 * f13 * f22 * f33 - d1 * f12 * f23 * f33 + d2 * f12 * f23 * f33 -
                d1 * f1 * f24 * f33 - d1 * f13 * f2 * f34 - d1 * f12 * f22 * f34 + 2 * d1 * f1 * f23 * f34 + d1 * f12 * f2 * f35 -
                d1 * f1 * f22 * f35 + 4 * f12 * f23 * f32 * v1 - 3 * f1 * f24 * f32 * v1 - 8 * f12 * f22 * f33 * v1 +
                4 * f1 * f23 * f33 * v1 + f24 * f33 * v1 + 4 * f12 * f2 * f34 * v1 + f1 * f22 * f34 * v1 - 2 * f23 * f34 * v1 -
                2 * f1 * f2 * f35 * v1 + f22 * f35 * v1 - f15 * f32 * v2 + 3 * f14 * f33 * v2 - 3 * f13 * f34 * v2 +
                f12 * f35 * v2 - f15 * f22 * v3 + 2 * f14 * f23 * v3 - f13 * f24 * v3 + 2 * f15 * f2 * f3 * v3 -
                f14 * f22 * f3 * v3 - 4 * f13 * f23 * f3 * v3 + 3 * f12 * f24 * f3 * v3 - 4 * f14 * f2 * f32 * v3 +
                8 * f13 * f22 * f32 * v3 - 4 * f12 * f23 * f32 * v3) /
               den);
    r.a_d1 = -((-(d2 * f15 * f22) + 2 * d2 * f14 * f23 - d2 * f13 * f24 - d2 * f14 * f22 * f3 + 2 * d1 * f13 * f23 * f3 +
                2 * d2 * f13 * f23 * f3 - 2 * d1 * f12 * f24 * f3 - d2 * f12 * f24 * f3 + d2 * f15 * f32 -
                3 * d1 * f13 * f22 * f32 - d2 * f13 * f22 * f32 + 2 * d1 * f12 * f23 * f32 - 2 * d2 * f12 * f23 * f32 +
                d1 * f1 * f24 * f32 + 2 * d2 * f1 * f24 * f32 - d2 * f14 * f33 + d1 * f12 * f22 * f33 +
                3 * d2 * f12 * f22 * f33 - 2 * d1 * f1 * f23 * f33 - 2 * d2 * f1 * f23 * f33 + d1 * f24 * f33 +
                d1 * f13 * f34 + d1 * f1 * f22 * f34 - 2 * d1 * f23 * f34 - d1 * f12 * f35 + d1 * f22 * f35 -
                8 * f12 * f23 * f3 * v1 + 6 * f1 * f24 * f3 * v1 + 12 * f12 * f22 * f32 * v1 - 8 * f1 * f23 * f32 * v1 -
                4 * f12 * f34 * v1 + 2 * f1 * f35 * v1 + 2 * f15 * f3 * v2 - 4 * f14 * f32 * v2 + 4 * f12 * f34 * v2 -
                2 * f1 * f35 * v2 - 2 * f15 * f3 * v3 + 8 * f12 * f23 * f3 * v3 - 6 * f1 * f24 * f3 * v3 +
                4 * f14 * f32 * v3 - 12 * f12 * f22 * f32 * v3 + 8 * f1 * f23 * f32 * v3) /
               den);
    r.a_d2 = -((d2 * f15 * f2 - d1 * f13 * f23 - 3 * d2 * f13 * f23 + d1 * f12 * f24 + 2 * d2 * f12 * f24 - d2 * f15 * f3 +
                d2 * f14 * f2 * f3 - d1 * f12 * f23 * f3 + d2 * f12 * f23 * f3 + d1 * f1 * f24 * f3 - d2 * f1 * f24 * f3 -
                d2 * f14 * f32 + 3 * d1 * f13 * f2 * f32 + d2 * f13 * f2 * f32 - d1 * f1 * f23 * f32 + d2 * f1 * f23 * f32 -
                2 * d1 * f24 * f32 - d2 * f24 * f32 - 2 * d1 * f13 * f33 + 2 * d2 * f13 * f33 - d1 * f12 * f2 * f33 -
                3 * d2 * f12 * f2 * f33 + 3 * d1 * f23 * f33 + d2 * f23 * f33 + d1 * f12 * f34 - d1 * f1 * f2 * f34 +
                d1 * f1 * f35 - d1 * f2 * f35 + 4 * f12 * f23 * v1 - 3 * f1 * f24 * v1 + 4 * f1 * f23 * f3 * v1 -
                3 * f24 * f3 * v1 - 12 * f12 * f2 * f32 * v1 + 4 * f23 * f32 * v1 + 8 * f12 * f33 * v1 - f1 * f34 * v1 -
                f35 * v1 - f15 * v2 - f14 * f3 * v2 + 8 * f13 * f32 * v2 - 8 * f12 * f33 * v2 + f1 * f34 * v2 + f35 * v2 +
                f15 * v3 - 4 * f12 * f23 * v3 + 3 * f1 * f24 * v3 + f14 * f3 * v3 - 4 * f1 * f23 * f3 * v3 +
                3 * f24 * f3 * v3 - 8 * f13 * f32 * v3 + 12 * f12 * f2 * f32 * v3 - 4 * f23 * f32 * v3) /
               den);
    r.a_d3 = -((-2 * d2 * f14 * f2 + d1 * f13 * f22 + 3 * d2 * f13 * f22 - d1 * f1 * f24 - d2 * f1 * f24 + 2 * d2 * f14 * f3 -
                2 * d1 * f13 * f2 * f3 - 2 * d2 * f13 * f2 * f3 + d1 * f12 * f22 * f3 - d2 * f12 * f22 * f3 + d1 * f24 * f3 +
                d2 * f24 * f3 + d1 * f13 * f32 - d2 * f13 * f32 - 2 * d1 * f12 * f2 * f32 + 2 * d2 * f12 * f2 * f32 +
                d1 * f1 * f22 * f32 - d2 * f1 * f22 * f32 + d1 * f12 * f33 - d2 * f12 * f33 + 2 * d1 * f1 * f2 * f33 +
                2 * d2 * f1 * f2 * f33 - 3 * d1 * f22 * f33 - d2 * f22 * f33 - 2 * d1 * f1 * f34 + 2 * d1 * f2 * f34 -
                4 * f12 * f22 * v1 + 2 * f24 * v1 + 8 * f12 * f2 * f3 * v1 - 4 * f1 * f22 * f3 * v1 - 4 * f12 * f32 * v1 +
                8 * f1 * f2 * f32 * v1 - 4 * f22 * f32 * v1 - 4 * f1 * f33 * v1 + 2 * f34 * v1 + 2 * f14 * v2 -
                4 * f13 * f3 * v2 + 4 * f1 * f33 * v2 - 2 * f34 * v2 - 2 * f14 * v3 + 4 * f12 * f22 * v3 - 2 * f24 * v3 +
                4 * f13 * f3 * v3 - 8 * f12 * f2 * f3 * v3 + 4 * f1 * f22 * f3 * v3 + 4 * f12 * f32 * v3 -
                8 * f1 * f2 * f32 * v3 + 4 * f22 * f32 * v3) /
               den);
    r.a_d4 = -((d2 * f13 * f2 - d1 * f12 * f22 - 2 * d2 * f12 * f22 + d1 * f1 * f23 + d2 * f1 * f23 - d2 * f13 * f3 +
                2 * d1 * f12 * f2 * f3 + d2 * f12 * f2 * f3 - d1 * f1 * f22 * f3 + d2 * f1 * f22 * f3 - d1 * f23 * f3 -
                d2 * f23 * f3 - d1 * f12 * f32 + d2 * f12 * f32 - d1 * f1 * f2 * f32 - 2 * d2 * f1 * f2 * f32 +
                2 * d1 * f22 * f32 + d2 * f22 * f32 + d1 * f1 * f33 - d1 * f2 * f33 + 3 * f1 * f22 * v1 - 2 * f23 * v1 -
                6 * f1 * f2 * f3 * v1 + 3 * f22 * f3 * v1 + 3 * f1 * f32 * v1 - f33 * v1 - f13 * v2 + 3 * f12 * f3 * v2 -
                3 * f1 * f32 * v2 + f33 * v2 + f13 * v3 - 3 * f1 * f22 * v3 + 2 * f23 * v3 - 3 * f12 * f3 * v3 +
                6 * f1 * f2 * f3 * v3 - 3 * f22 * f3 * v3) /
               den);
}

// HM frequency-map parameters (PhenomHMWaveform.cu:2107-2286)
template <class S>
__device__ inline void hm_map(S Rholm, S Mf_RD_lm, S Mf_RD_22, int mm, nd<S> flm, bool for_amp,
                              S& a, S& b, S& fi, S& fr)
{
    S Mf_1_22 = for_amp ? AMP_JOIN_INS : PHI_JOIN_INS;
    fi = Mf_1_22 / Rholm;
    fr = Mf_RD_lm;
    S Ai = 2.0 / mm, Bi = 0.0;
    S Trd = for_amp ? (fr - Mf_RD_lm + Mf_RD_22) : (Rholm * fr);
    S Ti = 2.0 * fi / mm;
    S Am = (Trd - Ti) / (fr - fi);
    S Bm = Ti - fi * Am;
    S Ar = 1.0, Br = 0.0;
    if (for_amp)
        Br = -Mf_RD_lm + Mf_RD_22;
    else
        Ar = Rholm;
    if (flm > fi) {
        if (flm > fr) { a = Ar; b = Br; }
        else { a = Am; b = Bm; }
    } else {
        a = Ai; b = Bi;
    }
}

// The fields that phm_setup's PART 1 (the (2,2) reference) writes, copied
// from one record to another (they depend on the binary alone).
template <class S>
__device__ inline void copy_reference(PhmRecT<S>& r, const PhmRecT<S>& w)
{
    r.Mf_ref = w.Mf_ref;
    r.phi0 = w.phi0;
    r.t0 = w.t0;
    r.a_fRD = w.a_fRD;
    r.a_fDM = w.a_fDM;
    r.a_g1 = w.a_g1;
    r.a_g2 = w.a_g2;
    r.a_g3 = w.a_g3;
    r.a_fmax = w.a_fmax;
    r.a_rho1 = w.a_rho1;
    r.a_rho2 = w.a_rho2;
    r.a_rho3 = w.a_rho3;
    r.a_d0 = w.a_d0;
    r.a_d1 = w.a_d1;
    r.a_d2 = w.a_d2;
    r.a_d3 = w.a_d3;
    r.a_d4 = w.a_d4;
    r.a_pf23 = w.a_pf23;
    r.a_pf1 = w.a_pf1;
    r.a_pf43 = w.a_pf43;
    r.a_pf53 = w.a_pf53;
    r.a_pf2 = w.a_pf2;
    r.a_amp0 = w.a_amp0;
}

// Per-(binary, mode) setup. Masses in solar masses (m1_SI/MSUN), distance in m.
// Mf_RD_lm/Mf_DM_lm: this mode's QNM frequencies; Mf_RD_D/Mf_DM_D: the PhenomD
// (2,2) ringdown frequencies from the QNM table (IMRPhenomHMCore, PhenomHMWaveform.cu:2671-2844).
// Returns false for an unsupported mode.
// PART -1: the whole record. The record has two independent dependency chains,
// which a caller can run in different wavefronts (they write disjoint fields):
// PART 1, the (2,2) reference: the PhenomD amplitude coefficients and Mf_ref,
// phi0, t0 from a PhenomD phase with unit ringdown ratios (set up in a local
// record); PART 0, everything else (the phase with this mode's ratios, the
// maps, the PN amplitude inputs, the scalings).
template <class S, int PART = -1>
__device__ inline bool phm_setup(PhmRecT<S>& r, nd<S> m1_SI, nd<S> m2_SI, nd<S> chi1_in, nd<S> chi2_in,
                                 nd<S> distance, int ell, int mm, nd<S> Mf_RD_lm, nd<S> Mf_DM_lm,
                                 nd<S> Mf_RD_D, nd<S> Mf_DM_D, bool phenomd)
{
    S m1 = m1_SI / MSUN_SI, m2 = m2_SI / MSUN_SI, chi1 = chi1_in, chi2 = chi2_in;
    enforce_primary(m1, m2, chi1, chi2);
    if constexpr (PART == 1) {
        PhmRecT<S> t;
        setup_phenomd(t, m1, m2, chi1, chi2, 1.0, 1.0, Mf_RD_D, Mf_DM_D, true);
        S ph22, dph22;
        phase_imr(t, t.a_fmax, ph22, &dph22);
        r.Mf_ref = t.a_fmax;
        r.phi0 = 0.5 * (ph22 + 0.0);
        r.t0 = dph22;
        r.a_fRD = t.a_fRD; r.a_fDM = t.a_fDM; r.a_g1 = t.a_g1; r.a_g2 = t.a_g2; r.a_g3 = t.a_g3;
        r.a_fmax = t.a_fmax; r.a_rho1 = t.a_rho1; r.a_rho2 = t.a_rho2; r.a_rho3 = t.a_rho3;
        r.a_d0 = t.a_d0; r.a_d1 = t.a_d1; r.a_d2 = t.a_d2; r.a_d3 = t.a_d3; r.a_d4 = t.a_d4;
        r.a_pf23 = t.a_pf23; r.a_pf1 = t.a_pf1; r.a_pf43 = t.a_pf43; r.a_pf53 = t.a_pf53; r.a_pf2 = t.a_pf2;
        r.a_amp0 = t.a_amp0;
        return true;
    }
    S finmass, finspin;
    final_state(m1, m2, chi1, chi2, finmass, finspin);
    S Mf_RD_22, Mf_DM_22;
    bool ok = true;
    if (!phenomd)
        ok = hm_ringdown(2, 2, finmass, finspin, Mf_RD_22, Mf_DM_22);
    else {
        Mf_RD_22 = Mf_RD_D;
        Mf_DM_22 = Mf_DM_D;
    }
    r.ell = phenomd ? 2 : ell;
    r.mm = phenomd ? 2 : mm;
    r.phenomd = phenomd ? 1 : 0;
    if (!phenomd && !((ell == 2 && mm == 2) || (ell == 2 && mm == 1) || (ell == 3 && mm == 3) ||
                      (ell == 3 && mm == 2) || (ell == 4 && mm == 4) || (ell == 4 && mm == 3)))
        ok = false;

    // (2,2) PhenomD: f_ref -> amplitude peak, phi0 and t0 there (PhenomHMWaveform.cu:2752-2784)
    if constexpr (PART == -1) {
        setup_phenomd(r, m1, m2, chi1, chi2, 1.0, 1.0, Mf_RD_D, Mf_DM_D, true);
        r.Mf_ref = r.a_fmax;
        S ph22, dph22;
        phase_imr(r, r.Mf_ref, ph22, &dph22);
        r.phi0 = 0.5 * (ph22 + 0.0);
        r.t0 = dph22;
    }
    const S Mtot = (m1_SI + m2_SI) / MSUN_SI;
    r.amp_scale = Mtot * MRSUN_SI * Mtot * MTSUN_SI / distance;
    r.M_sec = (m1 + m2) * MTSUN_SI;
    const double cs[7] = {0.0, PI_2, 0.0, -PI_2, PI, PI_2, 0.0};
    r.cshift = cs[r.mm < 0 ? 0 : (r.mm > 6 ? 0 : r.mm)];

    // PN amplitude-ratio inputs (IMRPhenomHMOnePointFiveSpinPN with the enforced masses/spins)
    {
        S Min = m1 + m2;
        S M1 = m1 / Min, M2 = m2 / Min;
        S M = M1 + M2;
        const S eta = M1 * M2 / (M * M);
        S c[5];
        pn_habs_coeffs(r.ell, r.mm, eta, S(sqrt(1.0 - 4 * eta)), S(0.5 * (chi1 + chi2)), S(0.5 * (chi1 - chi2)), c);
        r.pn_c1 = c[0];
        r.pn_c2 = c[1];
        r.pn_c3 = c[2];
        r.pn_c4 = c[3];
        r.pn_c4i = c[4];
        r.pn_vscale = M * 2.0 * PI / (S)r.mm;
        r.pn_v2ratio = cbrt(2.0 / (S)r.mm);
    }

    if (phenomd) {
        r.ai = 1.0; r.bi = 0.0; r.am = 1.0; r.bm = 0.0; r.ar = 1.0; r.br = 0.0;
        r.iai = 1.0; r.iam = 1.0; r.iar = 1.0;
        r.fi = 1e300; r.fr = 1e300; r.PhDBconst = 0.0; r.PhDCconst = 0.0; r.PhDBAterm = 0.0; r.phaseC = 0.0;
        r.amp_fi = 1e300; r.amp_fr = 1e300; r.amp_Ai = 1.0; r.amp_Am = 1.0; r.amp_Bm = 0.0; r.amp_Br = 0.0;
        return ok;
    }

    // this mode's ratios and the PhenomD phase re-joined with them
    const S Rholm = Mf_RD_22 / Mf_RD_lm;
    const S Taulm = Mf_DM_lm / Mf_DM_22;
    setup_phenomd(r, m1, m2, chi1, chi2, Rholm, Taulm, Mf_RD_D, Mf_DM_D, false);

    // phase map + constants (IMRPhenomHMPhasePreComp, PhenomHMWaveform.cu:2317-2369)
    const double shift = 0.0001;
    S fi, fr;
    hm_map(Rholm, Mf_RD_lm, Mf_RD_22, mm, shift, false, r.ai, r.bi, fi, fr);
    hm_map(Rholm, Mf_RD_lm, Mf_RD_22, mm, fi + shift, false, r.am, r.bm, fi, fr);
    hm_map(Rholm, Mf_RD_lm, Mf_RD_22, mm, fr + shift, false, r.ar, r.br, fi, fr);
    r.fi = fi;
    r.fr = fr;
    r.iai = rcp_pos(r.ai);
    r.iam = rcp_pos(r.am);
    r.iar = rcp_pos(r.ar);
    r.PhDBconst = phase_only(r, r.am * fi + r.bm) / r.am;
    r.PhDCconst = phase_only(r, r.ar * fr + r.br) / r.ar;
    r.PhDBAterm = phase_only(r, r.ai * fi + r.bi) / r.ai;
    r.phaseC = phase_only(r, r.am * r.fr + r.bm) / r.am - r.PhDBconst + r.PhDBAterm;

    // amplitude map (region boundaries and linear pieces)
    {
        S a, b, afi, afr;
        hm_map(Rholm, Mf_RD_lm, Mf_RD_22, mm, 0.0, true, a, b, afi, afr);
        r.amp_fi = afi;
        r.amp_fr = afr;
        r.amp_Ai = 2.0 / mm;
        S Trd = afr - Mf_RD_lm + Mf_RD_22;
        S Ti = 2.0 * afi / mm;
        r.amp_Am = (Trd - Ti) / (afr - afi);
        r.amp_Bm = Ti - afi * r.amp_Am;
        r.amp_Br = -Mf_RD_lm + Mf_RD_22;
    }
    return ok;
}

// Amplitude, phase and t(f) of one (binary, mode) at frequency f in Hz
// (calculate_modes / calculate_modes_phenomd, PhenomHMWaveform.cu:2461-2662).
// ad_out (optional): the PhenomD amplitude at the mapped frequency, before the
// mode's PN factor and the distance scaling (amp = (A_D * factor) * amp_scale).
template <class S>
__device__ inline void phm_eval(const PhmRecT<S>& r, double f, S& amp, S& phase, S& tf, S* ad_out = nullptr)
{
    // A PhenomD-only record (run_phenomd) carries identity maps (a = 1, b = 0,
    // breakpoints at 1e300) and l = m = 2, so it takes the same path: every map
    // and offset below is then exact, and the amplitude factor is 1.
    const S Mf_wf = f * r.M_sec;

    // amplitude
    S a, b;
    if (Mf_wf > r.amp_fi) {
        if (Mf_wf > r.amp_fr) { a = 1.0; b = r.amp_Br; }
        else { a = r.amp_Am; b = r.amp_Bm; }
    } else {
        a = r.amp_Ai; b = 0.0;
    }
    const S Mf_amp = a * Mf_wf + b;
    P6T<S> p = pow6(Mf_amp);
    S A = amp_imr(r, Mf_amp, p);
    if (ad_out) *ad_out = A;
    A *= hm_amp_factor(r, Mf_wf, Mf_amp);
    amp = A * r.amp_scale;

    // phase and t(f): one of three linear frequency maps, then one PhenomD phase
    // evaluation; phi = ph / a - c1 + c2 with c1 = c2 = 0 on the first piece
    S sa, isa, sb, c1, c2;
    if (!(Mf_wf > r.fi)) {
        sa = r.ai; isa = r.iai; sb = r.bi; c1 = 0.0; c2 = 0.0;
    } else if (!(Mf_wf > r.fr)) {
        sa = r.am; isa = r.iam; sb = r.bm; c1 = r.PhDBconst; c2 = r.PhDBAterm;
    } else {
        sa = r.ar; isa = r.iar; sb = r.br; c1 = r.PhDCconst; c2 = r.phaseC;
    }
    S ph, dph;
    // below both joins the amplitude and phase maps are the same (2/m) Mf
    // (PhenomHMWaveform.cu:2200-2212, :2107-2150): the powers of Mf_amp serve both
    const S Mf_ph = sa * Mf_wf + sb;
    phase_imr(r, Mf_ph, ph, &dph, &p, value(Mf_ph) == value(Mf_amp));
    const S phi = r.cshift + ((ph * isa - c1) + c2);
    phase = -r.t0 * (Mf_wf - r.Mf_ref) + (phi - (r.mm * r.phi0));
    const double INV_TWO_PI = 0.15915494309189535;
    tf = (r.M_sec * (dph * INV_TWO_PI)) - ((r.t0 * INV_TWO_PI) * r.M_sec);
}

}  // namespace phm
}  // namespace bbhx
