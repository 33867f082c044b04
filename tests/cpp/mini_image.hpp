// Minimal image type with the traverser/accessor interface the drop-in
// header relies on (the subset of vigra::BasicImage it uses), for testing
// include/enblend_amd/enblend.hpp without VIGRA.
#pragma once

#include <vector>

namespace mini {

template <class T>
struct RGB {
    T v[3];
    RGB() : v{T(), T(), T()} {}
    RGB(T r, T g, T b) : v{r, g, b} {}
    T operator[](int i) const { return v[i]; }
    T& operator[](int i) { return v[i]; }
};

template <class T>
struct Traverser {
    T* base;
    long stride;
    int x, y;
};

template <class T>
struct PixelAccessor {
    typedef T value_type;
    T operator()(const Traverser<T>& t) const { return t.base[(long)t.y * t.stride + t.x]; }
    void set(const T& v, const Traverser<T>& t) const { t.base[(long)t.y * t.stride + t.x] = v; }
};

template <class T>
class Image {
public:
    typedef T value_type;
    typedef T PixelType;
    typedef Traverser<T> traverser;
    typedef Traverser<T> const_traverser;
    typedef PixelAccessor<T> ConstAccessor;
    typedef PixelAccessor<T> Accessor;

    Image(int w, int h) : w_(w), h_(h), d_((size_t)w * h) {}
    int width() const { return w_; }
    int height() const { return h_; }
    traverser upperLeft() const { return traverser{const_cast<T*>(d_.data()), w_, 0, 0}; }
    traverser lowerRight() const { return traverser{const_cast<T*>(d_.data()), w_, w_, h_}; }
    Accessor accessor() const { return Accessor(); }
    T& operator()(int x, int y) { return d_[(size_t)y * w_ + x]; }
    const T& operator()(int x, int y) const { return d_[(size_t)y * w_ + x]; }

private:
    int w_, h_;
    std::vector<T> d_;
};

// Argument objects of the call sites (vigra::triple / vigra::pair and the
// srcImageRange ... destImage factories), found by argument-dependent lookup
// as VIGRA's are.
template <class A, class B, class C>
struct triple {
    A first;
    B second;
    C third;
};

template <class A, class B>
struct pair {
    A first;
    B second;
};

template <class T>
triple<Traverser<T>, Traverser<T>, PixelAccessor<T>> srcImageRange(const Image<T>& im)
{
    return {im.upperLeft(), im.lowerRight(), im.accessor()};
}

template <class T>
triple<Traverser<T>, Traverser<T>, PixelAccessor<T>> destImageRange(Image<T>& im)
{
    return {im.upperLeft(), im.lowerRight(), im.accessor()};
}

template <class T>
pair<Traverser<T>, PixelAccessor<T>> srcImage(const Image<T>& im) { return {im.upperLeft(), im.accessor()}; }

template <class T>
pair<Traverser<T>, PixelAccessor<T>> maskImage(const Image<T>& im) { return {im.upperLeft(), im.accessor()}; }

template <class T>
pair<Traverser<T>, PixelAccessor<T>> destImage(Image<T>& im) { return {im.upperLeft(), im.accessor()}; }

// The VIGRA algorithms the call sites use around the pyramid templates,
// restated as plain loops over whole images.
template <class T, class M>
void initImageIf(triple<Traverser<T>, Traverser<T>, PixelAccessor<T>> dest, pair<Traverser<M>, PixelAccessor<M>> mask,
                 T v)
{
    for (int y = dest.first.y; y < dest.second.y; ++y)
        for (int x = dest.first.x; x < dest.second.x; ++x) {
            Traverser<M> m = mask.first;
            m.x += x - dest.first.x;
            m.y += y - dest.first.y;
            if (mask.second(m)) {
                Traverser<T> d = dest.first;
                d.x = x;
                d.y = y;
                dest.third.set(v, d);
            }
        }
}

}  // namespace mini
