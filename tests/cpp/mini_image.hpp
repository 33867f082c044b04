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

}  // namespace mini
