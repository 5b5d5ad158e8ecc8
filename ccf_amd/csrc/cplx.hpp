// Complex arithmetic shared by host code and gfx950 kernels.
//
// Conventions follow weisluke/ccf include/complex.cuh: abs = hypot,
// arg = atan2, log = (log|z|, arg z), integer powers by repeated
// multiplication. Division differs on purpose: kernels divide through one
// real reciprocal of |b|^2 (1/b = conj(b) / |b|^2) instead of Smith's
// three divides (complex.cuh:218-280); arguments here are O(1e-3..1e3),
// far from the range where Smith's scaling matters.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace ccf
{

template <typename T>
struct cplx
{
	T re, im;
	__host__ __device__ constexpr cplx() : re(0), im(0) {}
	__host__ __device__ constexpr cplx(T r) : re(r), im(0) {}
	__host__ __device__ constexpr cplx(T r, T i) : re(r), im(i) {}
	template <typename U>
	__host__ __device__ explicit constexpr cplx(const cplx<U>& o) : re(static_cast<T>(o.re)), im(static_cast<T>(o.im)) {}

	__host__ __device__ cplx conj() const { return cplx(re, -im); }
	__host__ __device__ T norm2() const { return re * re + im * im; }
	__host__ __device__ T abs() const { return hypot(re, im); }
	__host__ __device__ T arg() const { return atan2(im, re); }
	__host__ __device__ cplx log() const { return cplx(::log(abs()), arg()); }
	__host__ __device__ cplx exp() const { T e = ::exp(re); return cplx(e * cos(im), e * sin(im)); }
	__host__ __device__ cplx operator-() const { return cplx(-re, -im); }
	__host__ __device__ cplx& operator+=(cplx o) { re += o.re; im += o.im; return *this; }
	__host__ __device__ cplx& operator-=(cplx o) { re -= o.re; im -= o.im; return *this; }
	__host__ __device__ cplx& operator*=(cplx o)
	{
		T r = re * o.re - im * o.im;
		im = re * o.im + im * o.re;
		re = r;
		return *this;
	}
	__host__ __device__ cplx& operator*=(T s) { re *= s; im *= s; return *this; }
	__host__ __device__ cplx& operator/=(T s) { re /= s; im /= s; return *this; }
	__host__ __device__ cplx inv() const
	{
		const T n = re * re + im * im;
#if defined(__HIP_DEVICE_COMPILE__)
		// device FP64: the reciprocal estimate refined by two Newton steps
		// (each squares the ~2^-23 relative error) instead of the IEEE
		// division sequence
		T d;
		if constexpr (sizeof(T) == 8)
		{
			d = __builtin_amdgcn_rcp(n);
			T e = __builtin_fma(-n, d, T(1));
			d = __builtin_fma(d, e, d);
			e = __builtin_fma(-n, d, T(1));
			d = __builtin_fma(d, e, d);
		}
		else
		{
			d = T(1) / n;
		}
#else
		const T d = T(1) / n;
#endif
		return cplx(re * d, -im * d);
	}
	__host__ __device__ cplx& operator/=(cplx o) { return *this *= o.inv(); }
	// z^n (n >= 0) by repeated squaring: 2 log2(n) products instead of n,
	// rounding that differs from pow's at the level of a few ulps per
	// product (the FMM translations of the tree use it)
	__host__ __device__ cplx ipow(int n) const
	{
		cplx r(1, 0), b = *this;
		for (; n > 0; n >>= 1)
		{
			if (n & 1) r *= b;
			b *= b;
		}
		return r;
	}
	__host__ __device__ cplx pow(int n) const
	{
		cplx r(1, 0);
		if (n > 0)
		{
			for (int i = 0; i < n; i++) r *= *this;
		}
		else if (n < 0)
		{
			for (int i = 0; i > n; i--) r *= *this;
			r = r.inv();
		}
		return r;
	}
};

template <typename T> __host__ __device__ inline cplx<T> operator+(cplx<T> a, cplx<T> b) { return cplx<T>(a.re + b.re, a.im + b.im); }
template <typename T> __host__ __device__ inline cplx<T> operator-(cplx<T> a, cplx<T> b) { return cplx<T>(a.re - b.re, a.im - b.im); }
template <typename T> __host__ __device__ inline cplx<T> operator*(cplx<T> a, cplx<T> b) { return cplx<T>(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
template <typename T> __host__ __device__ inline cplx<T> operator*(cplx<T> a, T s) { return cplx<T>(a.re * s, a.im * s); }
template <typename T> __host__ __device__ inline cplx<T> operator*(T s, cplx<T> a) { return cplx<T>(s * a.re, s * a.im); }
template <typename T> __host__ __device__ inline cplx<T> operator/(cplx<T> a, T s) { return cplx<T>(a.re / s, a.im / s); }
template <typename T> __host__ __device__ inline cplx<T> operator/(cplx<T> a, cplx<T> b) { return a * b.inv(); }
template <typename T> __host__ __device__ inline cplx<T> operator/(T s, cplx<T> b) { return s * b.inv(); }
template <typename T> __host__ __device__ inline cplx<T> operator+(cplx<T> a, T s) { return cplx<T>(a.re + s, a.im); }
template <typename T> __host__ __device__ inline cplx<T> operator-(cplx<T> a, T s) { return cplx<T>(a.re - s, a.im); }
template <typename T> __host__ __device__ inline cplx<T> operator+(T s, cplx<T> a) { return cplx<T>(s + a.re, a.im); }
template <typename T> __host__ __device__ inline cplx<T> operator-(T s, cplx<T> a) { return cplx<T>(s - a.re, -a.im); }

using cd = cplx<double>;
using cf = cplx<float>;

constexpr double PI = 3.141592653589793238462643383279502884;

} // namespace ccf
