// exp and sin for the ARAD kernels, rounding exactly like the host C
// library the reference's Fortran calls: shared by the device code and the
// host-side check (tests/exp_neg_check.cpp).
//
// An ARAD kernel entry at sigma = 0.1 amplifies one ulp of an overlap S by
// about 100 x 2 S / D (D = self1 + self2 - 2 S T cancels), to ~4e-13 of the
// entry: the 1e-12 parity target is met only where every exp and sin of the
// overlap rounds as the reference's do.  So these functions reproduce the
// algorithms of glibc's libm (2.28 and later, as built for x86-64 with FMA
// contraction; the host test compares them with the host libm):
//
// exp_libm(x) -- glibc's exp (sysdeps/ieee754/dbl-64/e_exp.c, the Arm
//   optimized-routines method, a Tang-style table-driven exp):
//   k = round(x * 128/ln2) (fused into x * 128/ln2 + 1.5 * 2^52), r = x -
//   k ln2/128 with a two-part ln2/128, 2^(j/128) = scale (1 + tail) from a
//   128-entry table, j = k mod 128, scaled by 2^floor(k/128) (exact: glibc
//   adds k << 45 to the scale's bits, here ldexp), e^r - 1 by a degree-5
//   polynomial, e^x = scale + scale (tail + r + r^2 (C2 + r C3) + r^4 (C4 +
//   r C5)).
//   |x| >= 512 takes the library's scaled path (products not fused, the
//   subnormal range rounded once), |x| >= 1024 gives 0 or overflow.
// exp_neg(x) -- the same for the ARAD overlap's exponents, x in (-16, 0]
//   ((8w)^2 * (-1/(4 w^2)) = -16): the main path only, no branches.
// sin_libm(y), 0 <= y < 2.426 -- glibc's sin (sysdeps/ieee754/dbl-64/
//   s_sin.c, IBM Accurate Mathematical Library): y rounded to the nearest
//   multiple of 1/128 (adding 1.5 * 2^45), sin and cos of that point from a
//   head + tail table, the remainder's sin and cos by short polynomials,
//   combined as sin(a + x) = sn + (ssn + s ccs - sn c + cs s) below 0.855469
//   and as cos(pi/2 - y) above; below 0.126 glibc uses a Taylor polynomial
//   (within 0.501 ulp) that is not reproduced: sin_cr, the correctly rounded
//   sin, stands in there (sin(y) = y below 2^-26).
// sin_cr(y), 0 <= y <= pi/2: double-double Taylor series, correctly rounded.
//
// The table entries are values of 2^(j/128), sin(i/128) and cos(i/128)
// rounded from 200-bit evaluations.
#pragma once

#include <math.h>
#include <stdint.h>

#ifndef FMLH_HD
#define FMLH_HD __host__ __device__
#endif

// Every function below starts with FMLH_NO_CONTRACT: its results depend on
// each operation being rounded as written (hipcc contracts a * b + c by
// default).  Host builds take -ffp-contract=off.
#if defined(__clang__)
#define FMLH_NO_CONTRACT _Pragma("clang fp contract(off)")
#else
#define FMLH_NO_CONTRACT
#endif

namespace fmlh {

// 2^(j/128) = scale (1 + tail), scale the correctly rounded 2^(j/128)
struct alignas(16) ExpPair {
  double tail, scale;
};

// sin(i/128) = sn + ssn, cos(i/128) = cs + ccs
struct SinCosEntry {
  double sn, ssn, cs, ccs;
};

}  // namespace fmlh

// {tail, scale}, j = 0..127
#define FMLH_EXP2_128_PAIRS \
    {0x0.0p+0, 0x1.0000000000000p+0}, {0x1.b3b4f1a88bf6ep-54, 0x1.0163da9fb3335p+0}, \
    {-0x1.160139cd8dc5dp-56, 0x1.02c9a3e778061p+0}, {-0x1.05e7a108766d1p-54, 0x1.04315e86e7f85p+0}, \
    {0x1.cd2523567f613p-55, 0x1.059b0d3158574p+0}, {-0x1.bce8023f98efap-55, 0x1.0706b29ddf6dep+0}, \
    {0x1.0f74e61e6c861p-57, 0x1.0874518759bc8p+0}, {0x1.0a3e45b33d399p-54, 0x1.09e3ecac6f383p+0}, \
    {0x1.79aa65d837b6dp-54, 0x1.0b5586cf9890fp+0}, {0x1.eb51a92fdeffcp-55, 0x1.0cc922b7247f7p+0}, \
    {0x1.ebe3d702f9cd1p-60, 0x1.0e3ec32d3d1a2p+0}, {-0x1.a033489906e0bp-57, 0x1.0fb66affed31bp+0}, \
    {-0x1.556522a2fbd0ep-54, 0x1.11301d0125b51p+0}, {-0x1.080ef8c4eea55p-58, 0x1.12abdc06c31ccp+0}, \
    {-0x1.1c923b9d5f416p-54, 0x1.1429aaea92de0p+0}, {0x1.0d3e3e95c55afp-55, 0x1.15a98c8a58e51p+0}, \
    {-0x1.01b15eaa59348p-55, 0x1.172b83c7d517bp+0}, {-0x1.f1ff055de323dp-55, 0x1.18af9388c8deap+0}, \
    {0x1.b898c3f1353bfp-55, 0x1.1a35beb6fcb75p+0}, {-0x1.6d99c7611eb26p-54, 0x1.1bbe084045cd4p+0}, \
    {0x1.aecf73e3a2f60p-54, 0x1.1d4873168b9aap+0}, {-0x1.fe782cb86389dp-55, 0x1.1ed5022fcd91dp+0}, \
    {0x1.a6f4144a6c38dp-55, 0x1.2063b88628cd6p+0}, {0x1.07a05b0e4047dp-55, 0x1.21f49917ddc96p+0}, \
    {0x1.68efde3a8a894p-54, 0x1.2387a6e756238p+0}, {0x1.75e18f274487dp-55, 0x1.251ce4fb2a63fp+0}, \
    {0x1.0472b981fe7f2p-55, 0x1.26b4565e27cddp+0}, {-0x1.6b87b3f71085ep-54, 0x1.284dfe1f56381p+0}, \
    {0x1.2f7e16d09ab31p-55, 0x1.29e9df51fdee1p+0}, {-0x1.d219b1a6fbffap-60, 0x1.2b87fd0dad990p+0}, \
    {0x1.b3782720c0ab4p-55, 0x1.2d285a6e4030bp+0}, {0x1.e149289cecb8fp-57, 0x1.2ecafa93e2f56p+0}, \
    {0x1.34d754db0abb6p-55, 0x1.306fe0a31b715p+0}, {0x1.64201e2ac744cp-55, 0x1.32170fc4cd831p+0}, \
    {0x1.fdd395dd3f84ap-55, 0x1.33c08b26416ffp+0}, {-0x1.6a3803b8e5b04p-55, 0x1.356c55f929ff1p+0}, \
    {-0x1.24aedcc4b5068p-54, 0x1.371a7373aa9cbp+0}, {-0x1.907f81b512d8ep-54, 0x1.38cae6d05d866p+0}, \
    {-0x1.1d1e83e9436d2p-56, 0x1.3a7db34e59ff7p+0}, {-0x1.91919b3ce1b15p-54, 0x1.3c32dc313a8e5p+0}, \
    {0x1.59f48a72a4c6dp-55, 0x1.3dea64c123422p+0}, {-0x1.312607a28698ap-54, 0x1.3fa4504ac801cp+0}, \
    {-0x1.8a78f4817895bp-58, 0x1.4160a21f72e2ap+0}, {-0x1.c2c9b67499a1bp-56, 0x1.431f5d950a897p+0}, \
    {0x1.363ed60c2ac11p-59, 0x1.44e086061892dp+0}, {0x1.666093b0664efp-54, 0x1.46a41ed1d0057p+0}, \
    {0x1.ecce1daa10379p-57, 0x1.486a2b5c13cd0p+0}, {0x1.3ff8e3f0f1230p-54, 0x1.4a32af0d7d3dep+0}, \
    {0x1.690cebb7aafb0p-56, 0x1.4bfdad5362a27p+0}, {0x1.31dbdeb54e077p-54, 0x1.4dcb299fddd0dp+0}, \
    {-0x1.f94340071a38ep-55, 0x1.4f9b2769d2ca7p+0}, {-0x1.7deccdc93a349p-55, 0x1.516daa2cf6642p+0}, \
    {-0x1.8dec6bd0f385fp-56, 0x1.5342b569d4f82p+0}, {-0x1.61246ec7b5cf6p-55, 0x1.551a4ca5d920fp+0}, \
    {0x1.3350518fdd78ep-54, 0x1.56f4736b527dap+0}, {0x1.b98b72f8a9b05p-56, 0x1.58d12d497c7fdp+0}, \
    {0x1.063e1e21c5409p-54, 0x1.5ab07dd485429p+0}, {0x1.4c7855019c6eap-60, 0x1.5c9268a5946b7p+0}, \
    {0x1.432e62b64c035p-54, 0x1.5e76f15ad2148p+0}, {-0x1.ce44a6199769fp-55, 0x1.605e1b976dc09p+0}, \
    {-0x1.c33c53bef4da8p-55, 0x1.6247eb03a5585p+0}, {-0x1.45378892be9aep-55, 0x1.6434634ccc320p+0}, \
    {-0x1.3cedd78565858p-54, 0x1.6623882552225p+0}, {0x1.710aa807e1964p-58, 0x1.68155d44ca973p+0}, \
    {-0x1.3b3efbf5e2228p-54, 0x1.6a09e667f3bcdp+0}, {-0x1.a12ad8734b982p-57, 0x1.6c012750bdabfp+0}, \
    {-0x1.367efb86da9eep-57, 0x1.6dfb23c651a2fp+0}, {-0x1.0dc3d54e08851p-55, 0x1.6ff7df9519484p+0}, \
    {-0x1.81f647e5a3ecfp-56, 0x1.71f75e8ec5f74p+0}, {-0x1.6ee4ac08b7db0p-55, 0x1.73f9a48a58174p+0}, \
    {-0x1.619321e55e68ap-55, 0x1.75feb564267c9p+0}, {0x1.09ccb5e09d4d3p-54, 0x1.780694fde5d3fp+0}, \
    {-0x1.b32dcb94da51dp-56, 0x1.7a11473eb0187p+0}, {0x1.4ecfd5467c06bp-54, 0x1.7c1ed0130c132p+0}, \
    {0x1.5ebe1abd66c55p-57, 0x1.7e2f336cf4e62p+0}, {-0x1.8a1c52fb3cf42p-55, 0x1.80427543e1a12p+0}, \
    {-0x1.369b6f13b3734p-54, 0x1.82589994cce13p+0}, {-0x1.05e843a19ff1ep-55, 0x1.8471a4623c7adp+0}, \
    {-0x1.4d450d872576ep-54, 0x1.868d99b4492edp+0}, {0x1.0ad675b0e8a00p-54, 0x1.88ac7d98a6699p+0}, \
    {0x1.db72fc1f0eab4p-55, 0x1.8ace5422aa0dbp+0}, {-0x1.5b6609cc5e7ffp-57, 0x1.8cf3216b5448cp+0}, \
    {0x1.bf68359f35f44p-56, 0x1.8f1ae99157736p+0}, {-0x1.3091fa71e3d83p-54, 0x1.9145b0b91ffc6p+0}, \
    {-0x1.da9b88b6c1e29p-58, 0x1.93737b0cdc5e5p+0}, {-0x1.c23f97c90b959p-57, 0x1.95a44cbc8520fp+0}, \
    {-0x1.2434322f4f9aap-54, 0x1.97d829fde4e50p+0}, {-0x1.5ca6cd7668e4bp-55, 0x1.9a0f170ca07bap+0}, \
    {0x1.1affc2b91ce27p-56, 0x1.9c49182a3f090p+0}, {0x1.dd235e10a73bbp-57, 0x1.9e86319e32323p+0}, \
    {-0x1.7c50422622263p-55, 0x1.a0c667b5de565p+0}, {0x1.b1c86e3e231d5p-55, 0x1.a309bec4a2d33p+0}, \
    {-0x1.1bbd1d3bcbb15p-54, 0x1.a5503b23e255dp+0}, {0x1.0cc319cee31d2p-54, 0x1.a799e1330b358p+0}, \
    {0x1.469846e735ab3p-55, 0x1.a9e6b5579fdbfp+0}, {-0x1.2dfcd978e9db4p-55, 0x1.ac36bbfd3f37ap+0}, \
    {0x1.c1a7792cb3387p-55, 0x1.ae89f995ad3adp+0}, {-0x1.07b8f4ad1d9fap-54, 0x1.b0e07298db666p+0}, \
    {-0x1.5c3d956dcaebap-58, 0x1.b33a2b84f15fbp+0}, {-0x1.0a40e3da6f640p-54, 0x1.b59728de5593ap+0}, \
    {-0x1.8d6f438ad9334p-57, 0x1.b7f76f2fb5e47p+0}, {-0x1.1eee26b588a35p-54, 0x1.ba5b030a1064ap+0}, \
    {0x1.4ffd70a5fddcdp-56, 0x1.bcc1e904bc1d2p+0}, {-0x1.1bdfbfa9298acp-54, 0x1.bf2c25bd71e09p+0}, \
    {0x1.36eae30af0cb3p-56, 0x1.c199bdd85529cp+0}, {0x1.ee3325c9ffd94p-55, 0x1.c40ab5fffd07ap+0}, \
    {0x1.4e08fd10959acp-55, 0x1.c67f12e57d14bp+0}, {0x1.3cdaf384e1a67p-57, 0x1.c8f6d9406e7b5p+0}, \
    {0x1.76b2c6c921968p-57, 0x1.cb720dcef9069p+0}, {-0x1.08a1883ccb5d2p-55, 0x1.cdf0b555dc3fap+0}, \
    {-0x1.fad5d3ffffa6fp-55, 0x1.d072d4a07897cp+0}, {-0x1.00dae3875a949p-54, 0x1.d2f87080d89f2p+0}, \
    {0x1.4a385a63d07a7p-56, 0x1.d5818dcfba487p+0}, {-0x1.2919e2040220fp-55, 0x1.d80e316c98398p+0}, \
    {0x1.e5a50d5c192acp-55, 0x1.da9e603db3285p+0}, {0x1.43a59ac016b4bp-55, 0x1.dd321f301b460p+0}, \
    {-0x1.2d52107b43e1fp-55, 0x1.dfc97337b9b5fp+0}, {-0x1.92ab93b470dc9p-55, 0x1.e264614f5a129p+0}, \
    {0x1.4b604603a88d3p-56, 0x1.e502ee78b3ff6p+0}, {0x1.3c5ec519d7271p-55, 0x1.e7a51fbc74c83p+0}, \
    {-0x1.ff7128fd391f0p-55, 0x1.ea4afa2a490dap+0}, {-0x1.dae98e223747dp-55, 0x1.ecf482d8e67f1p+0}, \
    {0x1.ec3bc41aa2008p-55, 0x1.efa1bee615a27p+0}, {0x1.42b94c3a9eb32p-55, 0x1.f252b376bba97p+0}, \
    {0x1.a64a931d185eep-55, 0x1.f50765b6e4540p+0}, {-0x1.e37bae43be3edp-55, 0x1.f7bfdad9cbe14p+0}, \
    {0x1.7893b4d91cd9dp-56, 0x1.fa7c1819e90d8p+0}, {0x1.305c14160cc89p-58, 0x1.fd3c22b8f71f1p+0}

// {sn, ssn, cs, ccs} at i/128, i = 0..115
#define FMLH_SINCOS_128 \
    {0x0.0p+0, 0x0.0p+0, 0x1.0000000000000p+0, 0x0.0p+0}, \
    {0x1.fffeaaaaeeeefp-8, -0x1.e45e2ec67b77cp-62, 0x1.fffc000155552p-1, 0x1.f4a01a0196daep-55}, \
    {0x1.fffaaaaeeeed5p-7, -0x1.2ab639a9f0776p-63, 0x1.fff000155549fp-1, 0x1.28a28a03a5ef3p-55}, \
    {0x1.7ff7001033255p-6, 0x1.efe2b51527336p-64, 0x1.ffdc006bff7e6p-1, 0x1.ae6dae86977bdp-55}, \
    {0x1.ffeaaaeeee86fp-6, -0x1.cd406fb224ae2p-60, 0x1.ffc00155527d3p-1, -0x1.3b54492d89b5bp-55}, \
    {0x1.3feb2b12d45d5p-5, 0x1.4ec54203d1c11p-60, 0x1.ff9c03414a7bap-1, 0x1.991f4be6c59bfp-57}, \
    {0x1.7fdc01032fba9p-5, -0x1.599bdf46e997ap-59, 0x1.ff7006bfdf99fp-1, -0x1.8b3b560648d5fp-56}, \
    {0x1.bfc6d78586dacp-5, 0x1.8e4fd03dbf236p-62, 0x1.ff3c0c8103a31p-1, 0x1.4856dbddc0e66p-56}, \
    {0x1.ffaaaeeed4edbp-5, -0x1.2d16d32684b69p-59, 0x1.ff0015549f4d3p-1, 0x1.328387b99426fp-55}, \
    {0x1.1fc343d808befp-4, -0x1.f3d32e6f3be4fp-58, 0x1.febc222a8ef9fp-1, 0x1.7934934f54c77p-58}, \
    {0x1.3facb12d1755bp-4, -0x1.921915299468bp-58, 0x1.fe7034129ef6fp-1, -0x1.cbf4337c96f97p-57}, \
    {0x1.5f911fd10b737p-4, -0x1.0184f02be9102p-58, 0x1.fe1c4c3c873ebp-1, -0x1.5a9c9057c4a02p-60}, \
    {0x1.7f701032550e4p-4, 0x1.afc2d1800501ap-60, 0x1.fdc06bf7e6b9bp-1, 0x1.31902b535f8dbp-55}, \
    {0x1.9f4902d55d1f9p-4, 0x1.2696d7eac1dc1p-58, 0x1.fd5c94b43e000p-1, -0x1.2e768cb4f92f9p-57}, \
    {0x1.bf1b78568391dp-4, 0x1.e91841dea4cc8p-58, 0x1.fcf0c800e99b1p-1, 0x1.ea3d786d186acp-57}, \
    {0x1.dee6f16c1cce6p-4, -0x1.50f8e2fb71673p-59, 0x1.fc7d078d1bc88p-1, 0x1.075d2447db685p-55}, \
    {0x1.feaaeee86ee36p-4, -0x1.afcb2bcc6f03bp-59, 0x1.fc015527d5bd3p-1, 0x1.b68f35094efb8p-55}, \
    {0x1.0f3378ddd71d1p-3, 0x1.d8468724f0f9ep-57, 0x1.fb7db2bfe0695p-1, 0x1.21dadf4f65ab1p-55}, \
    {0x1.1f0d3d7afceafp-3, -0x1.6ef95099769a5p-57, 0x1.faf22263c4bd3p-1, -0x1.52ace133a2769p-58}, \
    {0x1.2ee285e4ab88fp-3, -0x1.e4d0f05dee058p-57, 0x1.fa5ea641c36f2p-1, 0x1.04da6ed17cc7cp-59}, \
    {0x1.3eb312c5d66cbp-3, 0x1.47d666b66cb91p-57, 0x1.f9c340a7cc428p-1, 0x1.c5b6b063b7462p-55}, \
    {0x1.4e7ea4dc5f27bp-3, 0x1.949db2ac072fcp-58, 0x1.f91ff40374d01p-1, -0x1.7d03f4d3a9e4cp-57}, \
    {0x1.5e44fcfa126f3p-3, -0x1.6f443063f89b6p-57, 0x1.f874c2e1eecf6p-1, -0x1.c6514e1332b16p-55}, \
    {0x1.6e05dc05a4d4cp-3, -0x1.32c5c8b81c919p-66, 0x1.f7c1afeffde24p-1, -0x1.8f55bc47540b1p-56}, \
    {0x1.7dc102fbaf2b5p-3, 0x1.5ab50e23c97c3p-59, 0x1.f706bdf9ece1cp-1, -0x1.698c80c36dcb4p-55}, \
    {0x1.8d7632efaa944p-3, -0x1.20fa262cbb953p-57, 0x1.f643efeb82acdp-1, 0x1.6b00ac1fe28acp-56}, \
    {0x1.9d252d0cec312p-3, 0x1.9c43d80b1137dp-58, 0x1.f57948cff6797p-1, 0x1.e3a0d3e03b1d4p-57}, \
    {0x1.accdb297a0765p-3, -0x1.9883b57d6cdeap-58, 0x1.f4a6cbd1e3a79p-1, 0x1.13df0edaebb57p-55}, \
    {0x1.bc6f84edc6199p-3, 0x1.9c1a56a7b0cabp-57, 0x1.f3cc7c3b3d16ep-1, -0x1.21a3ad28a3494p-57}, \
    {0x1.cc0a6588289a3p-3, -0x1.868d09bc87c6bp-57, 0x1.f2ea5d753ffedp-1, 0x1.cc4215f56d583p-55}, \
    {0x1.db9e15fb5a5d0p-3, -0x1.32e20d6cc6fc2p-57, 0x1.f20073086649fp-1, 0x1.b940416c1984bp-56}, \
    {0x1.eb2a57f8ae5a3p-3, -0x1.0be06af572cebp-57, 0x1.f10ec09c5873bp-1, 0x1.d9072762c1283p-55}, \
    {0x1.faaeed4f31577p-3, -0x1.15d88508e32b8p-57, 0x1.f01549f7deea1p-1, 0x1.d3c1e99e5cafdp-55}, \
    {0x1.0515cbf65155cp-2, -0x1.9b8c29dfd8ec7p-56, 0x1.ef141300d2f26p-1, -0x1.2aa1b08ded372p-55}, \
    {0x1.0cd00cef36436p-2, -0x1.9fb0a0c93e2b4p-56, 0x1.ee0b1fbc0f11cp-1, -0x1.bfd2380bbc3b1p-59}, \
    {0x1.14861aa94ddebp-2, -0x1.be881b5b615a4p-57, 0x1.ecfa744d5efa1p-1, -0x1.56d0a4af541d0p-58}, \
    {0x1.1c37d64c6b876p-2, 0x1.46076fe0dcff4p-56, 0x1.ebe214f76efa8p-1, -0x1.02f9f12ba543ep-55}, \
    {0x1.23e52111aaf36p-2, -0x1.4f080334eff18p-56, 0x1.eac2061bbaf4fp-1, 0x1.2c1d53e94658dp-57}, \
    {0x1.2b8ddc43eb49fp-2, 0x1.1553899f2d807p-57, 0x1.e99a4c3a7cd83p-1, -0x1.2264b1bc53ce8p-55}, \
    {0x1.3331e94049f87p-2, 0x1.e0cb6b40c302cp-56, 0x1.e86aebf29a9edp-1, 0x1.9397afdbb58a7p-55}, \
    {0x1.3ad129769d3d8p-2, 0x1.03d550487839ap-63, 0x1.e733ea0193d40p-1, -0x1.6428b3546ce13p-55}, \
    {0x1.426b7e69ee697p-2, -0x1.f09c75705c59fp-56, 0x1.e5f54b436e9d0p-1, 0x1.7eb0fd02fc8bcp-55}, \
    {0x1.4a00c9b0f3d20p-2, 0x1.823ba6bb08eadp-56, 0x1.e4af14b2a449cp-1, -0x1.68ca02e8a6833p-55}, \
    {0x1.5190ecf68a77ap-2, 0x1.b357155eef0f3p-56, 0x1.e3614b680d6a5p-1, -0x1.27793aa015237p-56}, \
    {0x1.591bc9fa2f597p-2, 0x1.7c74bac3fe0cbp-57, 0x1.e20bf49acd6c1p-1, -0x1.660aec7ef636bp-58}, \
    {0x1.60a1429078775p-2, 0x1.b1fd80ba89133p-58, 0x1.e0af15a03dbcep-1, 0x1.fe8e702771ae6p-58}, \
    {0x1.682138a38d7f7p-2, -0x1.d889202444aadp-56, 0x1.df4ab3ebd875ep-1, -0x1.e2d8a7e6736c4p-55}, \
    {0x1.6f9b8e33a0255p-2, 0x1.42bc14ee9da0dp-56, 0x1.ddded50f228d6p-1, -0x1.e80c8d42ba2bfp-57}, \
    {0x1.7710255764214p-2, -0x1.6ead7314bb6cep-57, 0x1.dc6b7eb995912p-1, 0x1.4b364776dcd35p-58}, \
    {0x1.7e7ee03c86d4ep-2, -0x1.b63bcdabf5af2p-56, 0x1.daf0b6b888e83p-1, 0x1.a249e2b5e5ceap-55}, \
    {0x1.85e7a12826949p-2, 0x1.8a40e9b5face0p-56, 0x1.d96e82f71a9dcp-1, 0x1.ff61bd5d2039dp-55}, \
    {0x1.8d4a4a774992fp-2, 0x1.44a02ea766326p-56, 0x1.d7e4e97e17b4ap-1, -0x1.3b770352bed94p-57}, \
    {0x1.94a6be9f546c5p-2, -0x1.69ce13e683f58p-56, 0x1.d653f073e4040p-1, -0x1.76236434bec37p-55}, \
    {0x1.9bfce02e80510p-2, 0x1.09e39a320b0a4p-56, 0x1.d4bb9e1c619e0p-1, 0x1.f34bb77858f61p-55}, \
    {0x1.a34c91cc50ccap-2, -0x1.a310e3b50cecdp-58, 0x1.d31bf8d8d7c06p-1, 0x1.e60dd3089cbddp-56}, \
    {0x1.aa95b63a09277p-2, -0x1.6293eb13c0381p-57, 0x1.d1750727d94f0p-1, 0x1.0d52b1ec1a48ep-55}, \
    {0x1.b1d8305321617p-2, -0x1.ae242cb99f519p-56, 0x1.cfc6cfa52ad9fp-1, 0x1.8b5b5508f2a0dp-55}, \
    {0x1.b913e30dbac43p-2, -0x1.e38ad2f6c3ff1p-56, 0x1.ce115909a82e5p-1, 0x1.1f139bb31109ap-55}, \
    {0x1.c048b17b140a3p-2, 0x1.19fe6757e9fa7p-57, 0x1.cc54aa2b2972ep-1, 0x1.4ee162ba83a98p-57}, \
    {0x1.c7767ec7fd19ep-2, -0x1.eb14d1a3d5826p-58, 0x1.ca90c9fc67d0bp-1, -0x1.46a81485e3462p-57}, \
    {0x1.ce9d2e3d4a51fp-2, -0x1.2fc8a12dae298p-57, 0x1.c8c5bf8ce1a84p-1, 0x1.ab3d1a1590123p-56}, \
    {0x1.d5bca34047661p-2, 0x1.28a44a75fc29cp-56, 0x1.c6f39208be53bp-1, -0x1.741dbfbaadb42p-55}, \
    {0x1.dcd4c15329c9ap-2, 0x1.0d4c6e171fd9ap-56, 0x1.c51a48b8b175ep-1, -0x1.1bbb43b9aa880p-57}, \
    {0x1.e3e56c1582a69p-2, -0x1.0a4821099f88fp-58, 0x1.c339eb01ddd81p-1, -0x1.caaf5ee82c5c0p-55}, \
    {0x1.eaee8744b05f0p-2, -0x1.789b43c9b027dp-58, 0x1.c1528065b7d50p-1, -0x1.892111312e828p-55}, \
    {0x1.f1eff6bc4f97bp-2, 0x1.17212f8a7525cp-56, 0x1.bf641081e7536p-1, 0x1.b7bd71628a9a1p-55}, \
    {0x1.f8e99e76abc97p-2, 0x1.9d950af2d00a3p-58, 0x1.bd6ea310294f5p-1, 0x1.31bbcc88c109dp-56}, \
    {0x1.ffdb628d2f57ap-2, 0x1.f4a992e905b6ap-57, 0x1.bb723fe630f32p-1, 0x1.72bd2452d0a39p-56}, \
    {0x1.0362939c69955p-1, -0x1.2d8cd78397b01p-55, 0x1.b96eeef58840ep-1, 0x1.45a3cc78fade0p-58}, \
    {0x1.06d3686946e5bp-1, 0x1.3f5ae4538ff1bp-55, 0x1.b764b84b704c2p-1, -0x1.f5848c21b389bp-55}, \
    {0x1.0a4021e9e1001p-1, -0x1.6f643a13914f6p-55, 0x1.b553a410c104ep-1, 0x1.8ff7947027a15p-58}, \
    {0x1.0da8b26b5672ep-1, -0x1.a58def0bee909p-55, 0x1.b33bba89c8948p-1, 0x1.ea6a51d1f6ca9p-55}, \
    {0x1.110d0c4b69c3bp-1, 0x1.d918998809981p-55, 0x1.b11d04162a4c6p-1, 0x1.1dd561efbc0c2p-56}, \
    {0x1.146d21f8b7f82p-1, 0x1.bf9535e2739a8p-56, 0x1.aef78930bd275p-1, -0x1.f836279746f94p-56}, \
    {0x1.17c8e5f2eedb0p-1, 0x1.35e57102e2488p-57, 0x1.accb526f69de5p-1, 0x1.8fb6a8dd6b6ccp-55}, \
    {0x1.1b204acb02fddp-1, -0x1.f190c70cbb5fep-58, 0x1.aa98688308913p-1, -0x1.b83d607cd5072p-63}, \
    {0x1.1e7343236574cp-1, 0x1.22a3fa4f41d5ap-56, 0x1.a85ed4373e02dp-1, 0x1.9be06385ec792p-57}, \
    {0x1.21c1c1b0394cfp-1, 0x1.e5b324b23aa31p-58, 0x1.a61e9e72586afp-1, 0x1.58330e2fd453fp-55}, \
    {0x1.250bb93788bbbp-1, 0x1.ea3d02457bccep-56, 0x1.a3d7d0352bdcfp-1, -0x1.68dbaeca19669p-55}, \
    {0x1.28511c917a067p-1, -0x1.01df1d9a16b70p-55, 0x1.a18a729aee445p-1, 0x1.95e25736c0357p-60}, \
    {0x1.2b91dea88421ep-1, -0x1.fa371db216ab0p-55, 0x1.9f368ed912f85p-1, -0x1.1d200c5791606p-55}, \
    {0x1.2ecdf279a3082p-1, 0x1.d3557e0e7e37ep-55, 0x1.9cdc2e3f25e5cp-1, 0x1.3f99112993f62p-55}, \
    {0x1.32054b148bc4fp-1, 0x1.f6b42095a135bp-55, 0x1.9a7b5a36a6514p-1, 0x1.722cfcc9fa7a9p-55}, \
    {0x1.3537db9be0367p-1, 0x1.b327e7af040f0p-57, 0x1.98141c42e1310p-1, 0x1.d1ff80488f08dp-55}, \
    {0x1.386597456282bp-1, -0x1.10fada93b07a8p-56, 0x1.95a67e00cb1fdp-1, -0x1.0befda21f862dp-55}, \
    {0x1.3b8e715a2840ap-1, -0x1.97653a7d2f07ap-56, 0x1.93328926d9e92p-1, -0x1.bb77003600cdap-55}, \
    {0x1.3eb25d36cd53ap-1, -0x1.be570e1570fc0p-58, 0x1.90b84784ddaf7p-1, -0x1.0feb10ab93b87p-56}, \
    {0x1.41d14e4ba6790p-1, 0x1.4608fd287ecf5p-55, 0x1.8e37c303d9ad1p-1, -0x1.463a4b53d4bf8p-57}, \
    {0x1.44eb381cf386bp-1, -0x1.3ed6c1e6a5505p-55, 0x1.8bb105a5dc900p-1, 0x1.863e03e9474c1p-55}, \
    {0x1.48000e431159fp-1, -0x1.b194a7463ed10p-55, 0x1.89241985d871fp-1, 0x1.c48d9c413ed84p-55}, \
    {0x1.4b0fc46aab761p-1, 0x1.0da05738cc59cp-61, 0x1.869108d77a6c6p-1, 0x1.338ffe2bfe9ddp-56}, \
    {0x1.4e1a4e54ed51bp-1, -0x1.a492f89b7c76ap-55, 0x1.83f7dde701ca0p-1, -0x1.152cf609bc6e8p-59}, \
    {0x1.511f9fd7b351cp-1, -0x1.5c0e861c48831p-55, 0x1.8158a31916d5dp-1, -0x1.de8b90b8228dep-57}, \
    {0x1.541facddbb724p-1, 0x1.232c28520d391p-56, 0x1.7eb362eaa1488p-1, 0x1.a1d65a4a5959fp-58}, \
    {0x1.571a6966d59b3p-1, 0x1.c843b4d0fb197p-58, 0x1.7c0827f09e54fp-1, -0x1.c73d6d72aee68p-57}, \
    {0x1.5a0fc98813a12p-1, -0x1.d82e2b7d4227bp-55, 0x1.7956fcd7f6543p-1, -0x1.ab276e9d45ae4p-55}, \
    {0x1.5cffc16bf8f0dp-1, 0x1.96cb370eb578ap-55, 0x1.769fec655211fp-1, -0x1.827d5cf8c68c5p-57}, \
    {0x1.5fea4552a9e57p-1, 0x1.0b6cef7ee20b7p-55, 0x1.73e30174efba1p-1, -0x1.5d3ae3d94ad5fp-57}, \
    {0x1.62cf49921ac79p-1, -0x1.edd9855b6241ap-55, 0x1.712046fa77678p-1, 0x1.425b0a5029c81p-55}, \
    {0x1.65aec2963e755p-1, 0x1.126f96b71053cp-55, 0x1.6e57c800cf55ep-1, 0x1.60286dedbd0a6p-55}, \
    {0x1.6888a4e134b2fp-1, -0x1.6b7d37644d5e6p-55, 0x1.6b898fa9efb5dp-1, 0x1.15ac786ccf4b2p-56}, \
    {0x1.6b5ce50b7821ap-1, -0x1.5d5158f702e0fp-57, 0x1.68b5a92eb6253p-1, -0x1.9a91ad985f89cp-55}, \
    {0x1.6e2b77c40bde1p-1, -0x1.0e729857fad53p-56, 0x1.65dc1fdeb8cbap-1, -0x1.97c1b47337c77p-58}, \
    {0x1.70f451d0a8c40p-1, 0x1.97ede3885770dp-57, 0x1.62fcff20191c7p-1, 0x1.d9143895756efp-57}, \
    {0x1.73b7680dea578p-1, -0x1.2248306dc12a2p-56, 0x1.6018526f563dfp-1, 0x1.46ca5e0e432d0p-55}, \
    {0x1.7674af6f7b524p-1, 0x1.e9d3f94ac84a8p-56, 0x1.5d2e255f1f17ap-1, 0x1.0314104c8892bp-55}, \
    {0x1.792c1d0041d52p-1, -0x1.abf05eeb354ebp-55, 0x1.5a3e839824077p-1, 0x1.428aa2759be62p-55}, \
    {0x1.7bdda5e28b3c2p-1, 0x1.ad1197ccd0392p-59, 0x1.574978d8e83f2p-1, 0x1.f4714af282d23p-55}, \
    {0x1.7e893f5037959p-1, 0x1.0eefbaa650c4cp-55, 0x1.544f10f592ca5p-1, -0x1.e7ae8e6c7a62fp-55}, \
    {0x1.812ede9ae4ba4p-1, -0x1.7830adf402ddap-55, 0x1.514f57d7bf3dap-1, 0x1.47a108073c259p-56}, \
    {0x1.83ce792c1906ep-1, -0x1.f3899682b4a7dp-56, 0x1.4e4a597e4e10ep-1, 0x1.ccd992849f6c8p-56}, \
    {0x1.866804856db62p-1, 0x1.407b4e7476623p-57, 0x1.4b4021fd34a33p-1, -0x1.ee903cecc18cbp-55}, \
    {0x1.88fb7640b8da2p-1, -0x1.49987c11efaa3p-55, 0x1.4830bd7d4ceb3p-1, 0x1.df77ff20d5448p-55}, \
    {0x1.8b88c41036f4ep-1, -0x1.d1be754db84c3p-57, 0x1.451c383c24d43p-1, -0x1.54b060ea42adep-55}, \
    {0x1.8e0fe3beb42f8p-1, 0x1.324c55de9ed0bp-55, 0x1.42029e8bcd474p-1, 0x1.995705e2a2526p-55}, \
    {0x1.9090cb2fb52efp-1, -0x1.ca9146d0ee90bp-57, 0x1.3ee3fcd2a8e3cp-1, 0x1.c99ab27a9e083p-55}

namespace fmlh {

FMLH_HD inline double bits_to_double(uint64_t u) {
  double d;
  __builtin_memcpy(&d, &u, sizeof(d));
  return d;
}

FMLH_HD inline uint64_t double_to_bits(double d) {
  uint64_t u;
  __builtin_memcpy(&u, &d, sizeof(u));
  return u;
}

// Shared steps of exp_libm: the table scale times 2^(e + floor(k/128)) and
// the correction tmp, e^x = scale (1 + tmp).  c2 and c4 are the polynomial's
// constant terms, passed in so that device code can keep them in vector
// registers (an FMA reads at most one scalar register).
FMLH_HD inline double exp_reduce(double x, const ExpPair *__restrict__ tab, double &scale, int e = 0,
                                 double c2 = 0x1.ffffffffffdbdp-2, double c4 = 0x1.55555cf172b91p-5) {
  FMLH_NO_CONTRACT
  double kd = fma(0x1.71547652b82fep+7, x, 0x1.8p52);  // x * 128/ln2 rounded to an integer
  const int ki = (int)(uint32_t)double_to_bits(kd);   // the low word: k
  kd = kd - 0x1.8p52;
  const double r = fma(kd, -0x1.cf79abc9e3b3ap-47, fma(kd, -0x1.62e42fefa0000p-8, x));
  const ExpPair t = tab[ki & 127];
  scale = ldexp(t.scale, (ki >> 7) + e);
  const double r2 = r * r;
  const double p1 = fma(r, 0x1.555555555543cp-3, c2);  // C2 + r C3
  const double p2 = fma(r, 0x1.1111167a4d017p-7, c4);  // C4 + r C5
  return fma(r2 * r2, p2, fma(r2, p1, t.tail + r));
}

// x in (-16, 0] (any |x| < 512): the main path.
FMLH_HD inline double exp_neg(double x, const ExpPair *__restrict__ tab,
                              double c2 = 0x1.ffffffffffdbdp-2, double c4 = 0x1.55555cf172b91p-5) {
  FMLH_NO_CONTRACT
  double scale;
  const double tmp = exp_reduce(x, tab, scale, 0, c2, c4);
  return fma(scale, tmp, scale);
}

// exp_neg's polynomial constants C2 and C4.
#define FMLH_EXP_C2 0x1.ffffffffffdbdp-2
#define FMLH_EXP_C4 0x1.55555cf172b91p-5

// Any x.
FMLH_HD inline double exp_libm(double x, const ExpPair *__restrict__ tab) {
  FMLH_NO_CONTRACT
  const double ax = fabs(x);
  if (ax < 512.0) return exp_neg(x, tab);
  if (!(ax < 1024.0)) return x != x ? x : (x > 0.0 ? HUGE_VAL : 0.0);
  double scale;
  if (x > 0.0) {  // the scale's exponent may overflow: scale down by 2^1009
    const double tmp = exp_reduce(x, tab, scale, -1009);
    const double st = scale * tmp;
    return 0x1p1009 * (scale + st);
  }
  // k < 0: scale up by 2^1022, round once to the final precision in the
  // subnormal range, then scale back
  const double tmp = exp_reduce(x, tab, scale, 1022);
  const double st = scale * tmp;
  double y = scale + st;
  if (y < 1.0) {
    double lo = scale - y + st;
    const double hi = 1.0 + y;
    lo = 1.0 - hi + y + lo;
    y = (hi + lo) - 1.0;
    if (y == 0.0) y = 0.0;
  }
  return 0x1p-1022 * y;
}

// The name the atom-pair kernels use for their exponentials D * (-1/sigma^2).
FMLH_HD inline double exp_neg_wide(double x, const ExpPair *__restrict__ tab) {
  return exp_libm(x, tab);
}

// ---- double-double helpers for sin_cr ----

struct DD {
  double hi, lo;
};

FMLH_HD inline DD dd_two_sum(double a, double b) {
  FMLH_NO_CONTRACT
  const double s = a + b;
  const double bb = s - a;
  return DD{s, (a - (s - bb)) + (b - bb)};
}

FMLH_HD inline DD dd_fast_two_sum(double a, double b) {  // |a| >= |b|
  FMLH_NO_CONTRACT
  const double s = a + b;
  return DD{s, b - (s - a)};
}

FMLH_HD inline DD dd_mul(DD x, DD y) {
  FMLH_NO_CONTRACT
  const double p = x.hi * y.hi;
  double e = fma(x.hi, y.hi, -p);
  e = fma(x.hi, y.lo, e);
  e = fma(x.lo, y.hi, e);
  return dd_fast_two_sum(p, e);
}

FMLH_HD inline DD dd_add(DD x, DD y) {
  FMLH_NO_CONTRACT
  const DD s = dd_two_sum(x.hi, y.hi);
  return dd_fast_two_sum(s.hi, s.lo + (x.lo + y.lo));
}

// (-1)^k / (2k+1)!, k = 1..10, as double-doubles
#define FMLH_SIN_COEFS                                                                       \
  {-0x1.5555555555555p-3, -0x1.5555555555555p-57}, {0x1.1111111111111p-7, 0x1.1111111111111p-63}, \
  {-0x1.a01a01a01a01ap-13, -0x1.a01a01a01a01ap-73}, {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73}, \
  {-0x1.ae64567f544e4p-26, 0x1.c062e06d1f209p-80}, {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87}, \
  {-0x1.ae7f3e733b81fp-41, -0x1.1d8656b0ee8cbp-97}, {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103}, \
  {-0x1.2f49b46814157p-57, -0x1.2650f61dbdcb4p-112}, {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120}
// (-1)^k / (2k)!, k = 1..11
#define FMLH_COS_COEFS                                                                        \
  {-0x1.0000000000000p-1, 0.0}, {0x1.5555555555555p-5, 0x1.5555555555555p-59},                \
  {-0x1.6c16c16c16c17p-10, 0x1.f49f49f49f49fp-65}, {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76}, \
  {-0x1.27e4fb7789f5cp-22, -0x1.cbbc05b4fa99ap-76}, {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83}, \
  {-0x1.93974a8c07c9dp-37, -0x1.05d6f8a2efd1fp-92}, {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101}, \
  {-0x1.6827863b97d97p-53, -0x1.eec01221a8b0bp-107}, {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120}, \
  {-0x1.0ce396db7f853p-70, 0x1.aebcdbd20331cp-124}

// sin(y) correctly rounded for 0 <= y <= pi/2 + 1e-15.
FMLH_HD inline double sin_cr(double y) {
  FMLH_NO_CONTRACT
  const DD sc[10] = {FMLH_SIN_COEFS};
  const DD cc[11] = {FMLH_COS_COEFS};
  if (y <= 0x1.921fb54442d18p-1) {  // pi/4: y + y u S(u), u = y^2
    const DD u = dd_mul(DD{y, 0.0}, DD{y, 0.0});
    DD s = sc[9];
    for (int k = 8; k >= 0; --k) s = dd_add(dd_mul(s, u), sc[k]);
    const DD t = dd_mul(dd_mul(s, u), DD{y, 0.0});
    const DD r = dd_add(DD{y, 0.0}, t);
    return r.hi + r.lo;
  }
  // z = pi/2 - y (the head difference is exact: y >= (pi/2)/2), cos z
  const DD z = dd_two_sum(0x1.921fb54442d18p+0 - y, 0x1.1a62633145c07p-54);
  const DD u = dd_mul(z, z);
  DD c = cc[10];
  for (int k = 9; k >= 0; --k) c = dd_add(dd_mul(c, u), cc[k]);
  const DD r = dd_add(DD{1.0, 0.0}, dd_mul(c, u));
  return r.hi + r.lo;
}

// sin of a + x from the table point a = i/128 (glibc's do_sin with dx = 0).
FMLH_HD inline double sin_near(double y, const SinCosEntry *__restrict__ tab) {
  FMLH_NO_CONTRACT
  const double big = 0x1.8p45;
  const double u = big + y;
  const SinCosEntry e = tab[(uint32_t)double_to_bits(u)];
  const double x = y - (u - big);
  const double xx = x * x;
  const double s = x + fma(x * xx, fma(xx, 8.33333214285722277379541354343671e-03,
                                       -1.66666666666664880952546298448555e-01), 0.0);
  const double c = xx * fma(xx, fma(xx, 1.38888874007937613028114285595617e-03,
                                    -4.16666666666664434524222570944589e-02),
                            4.99999999999999999999950396842453e-01);
  const double cor = fma(e.cs, s, fma(-e.sn, c, fma(s, e.ccs, e.ssn)));
  return e.sn + cor;
}

// cos of x + dx (|x| < 0.72) from the table point nearest |x| (glibc's do_cos).
FMLH_HD inline double cos_near(double x, double dx, const SinCosEntry *__restrict__ tab) {
  FMLH_NO_CONTRACT
  const double big = 0x1.8p45;
  if (x < 0.0) dx = -dx;
  const double u = big + fabs(x);
  const SinCosEntry e = tab[(uint32_t)double_to_bits(u)];
  x = fabs(x) - (u - big) + dx;
  const double xx = x * x;
  const double s = fma(x * xx, fma(xx, 8.33333214285722277379541354343671e-03,
                                   -1.66666666666664880952546298448555e-01), x);
  const double c = xx * fma(xx, fma(xx, 1.38888874007937613028114285595617e-03,
                                    -4.16666666666664434524222570944589e-02),
                            4.99999999999999999999950396842453e-01);
  const double cor = fma(-e.sn, s, fma(-e.cs, c, fma(-s, e.ssn, e.ccs)));
  return e.cs + cor;
}

// 0 <= y < 2.426265 (y = d pi/(2 cut) for d <= cut).
FMLH_HD inline double sin_libm(double y, const SinCosEntry *__restrict__ tab) {
  FMLH_NO_CONTRACT
  if (y < 0.126) return y < 0x1p-26 ? y : sin_cr(y);
  if ((uint32_t)(double_to_bits(y) >> 32) < 0x3feb6000u) return sin_near(y, tab);  // < 0.855469
  return cos_near(0x1.921fb54442d18p+0 - y, 0x1.1a62633145c07p-54, tab);
}

}  // namespace fmlh
