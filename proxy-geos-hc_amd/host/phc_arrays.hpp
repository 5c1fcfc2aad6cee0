// phc_arrays.hpp -- host containers with the interface of the reference's
// USE_VECTOR containers (src/utils/dataType.hpp:14-80: Vector / Array2D with
// operator() and operator[] element access; allocateVector / allocateArray2D),
// plus the device binding of a wavefield array.
//
// A bound Array2D keeps one column per wavefield level (pnGlobal(node, level))
// and behaves like the reference's CHAI-managed arrays (dataType.hpp,
// USE_LVARRAY): the library holds the levels on the GPU; the host copy of a
// column is refreshed when it is read after a device step, and a column the
// host has written through a non-const access is written back to the device
// before the next device step.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <vector>

template <class T> class Vector {
  public:
    Vector() = default;
    explicit Vector(int n) : data_((size_t)(n > 0 ? n : 0)) {}
    T &operator()(int i) { return data_[(size_t)i]; }
    const T &operator()(int i) const { return data_[(size_t)i]; }
    T &operator[](int i) { return data_[(size_t)i]; }
    const T &operator[](int i) const { return data_[(size_t)i]; }
    int size() const { return (int)data_.size(); }
    T *data() { return data_.data(); }
    const T *data() const { return data_.data(); }
    void assign(size_t n, const T &v) { data_.assign(n, v); }

  private:
    std::vector<T> data_;
};

// Device side of a bound Array2D<T>: moves one column (rows values, in row
// order) between the host copy and the library's device level.
template <class T> struct ColumnStore {
    // cleared when the store goes away, so a bound array never calls into a
    // destroyed store
    std::shared_ptr<bool> alive = std::make_shared<bool>(true);
    ColumnStore() = default;
    ColumnStore(const ColumnStore &) = delete;
    ColumnStore &operator=(const ColumnStore &) = delete;
    virtual ~ColumnStore() { *alive = false; }
    virtual void pull(int col, T *host) = 0;
    virtual void push(int col, const T *host) = 0;
    // One element without moving the column; false if not supported.
    virtual bool read_one(int64_t row, int col, T *v) { (void)row; (void)col; (void)v; return false; }
};

template <class T> class Array2D {
  public:
    Array2D() = default;
    Array2D(int rows, int cols, T fill = T(0)) : rows_(rows > 0 ? rows : 0), cols_(cols > 0 ? cols : 0), fill_(fill)
    {
        state_.assign((size_t)cols_, Synced);
    }
    // A copy is a plain host array: the device binding stays with the original.
    Array2D(const Array2D &o) : rows_(o.rows_), cols_(o.cols_), fill_(o.fill_)
    {
        o.pull_stale();
        data_ = o.data_;
        state_.assign((size_t)cols_, Synced);
    }
    Array2D &operator=(const Array2D &o)
    {
        if (this != &o) {
            if (live()) throw std::runtime_error("assignment to an array whose levels live on the device");
            o.pull_stale();
            rows_ = o.rows_; cols_ = o.cols_; fill_ = o.fill_;
            data_ = o.data_;
            state_.assign((size_t)cols_, Synced);
        }
        return *this;
    }

    // Element access.  A non-const access may write: on a bound array it
    // marks the column as changed on the host.
    T &operator()(int64_t r, int c)
    {
        touch(c, true);
        return data_[(size_t)(r * cols_ + c)];
    }
    const T &operator()(int64_t r, int c) const
    {
        touch(c, false);
        return data_[(size_t)(r * cols_ + c)];
    }
    // Read one element; on a bound array whose column is newer on the device
    // this reads the single value from the device when the store supports it.
    T value(int64_t r, int c) const
    {
        if (live() && state_[(size_t)c] == DeviceNewer) {
            T v;
            if (store_->read_one(r, c, &v)) return v;
        }
        return (*this)(r, c);
    }
    int64_t rows() const { return rows_; }
    int cols() const { return cols_; }
    // Value every element holds until written (the host copy is allocated at
    // the first element access).
    T fill_value() const { return fill_; }
    void fill(T v)
    {
        fill_ = v;
        if (!data_.empty()) data_.assign(data_.size(), v);
        if (live()) {
            alloc();
            for (auto &s : state_) s = HostNewer;
        }
    }
    bool host_allocated() const { return !data_.empty(); }

    // ---- device binding (used by the solver classes)
    bool bound() const { return live(); }
    // The device takes the host's values: columns the host has allocated are
    // written to the device; an untouched array must equal the device's
    // initial state (its fill value).
    void bind(ColumnStore<T> *s)
    {
        store_ = s;
        alive_ = s->alive;
        state_.assign((size_t)cols_, host_allocated() ? HostNewer : Synced);
        push_host_changes();
    }
    void unbind()
    {
        pull_stale();
        store_ = nullptr;
        alive_.reset();
        state_.assign((size_t)cols_, Synced);
    }
    // Before a device step: write back the columns the host changed.
    void push_host_changes()
    {
        if (!live()) return;
        std::vector<T> buf;
        for (int c = 0; c < cols_; ++c)
            if (state_[(size_t)c] == HostNewer) {
                buf.resize((size_t)rows_);
                for (int64_t r = 0; r < rows_; ++r) buf[(size_t)r] = data_[(size_t)(r * cols_ + c)];
                store_->push(c, buf.data());
                state_[(size_t)c] = Synced;
            }
    }
    // After a device step: the device's column c is newer than the host copy.
    void device_changed(int c)
    {
        if (live()) state_[(size_t)c] = DeviceNewer;
    }

  private:
    enum State : uint8_t { Synced, DeviceNewer, HostNewer };
    // Bound to a store that still exists.  A column still newer on a store
    // that is gone has lost its values: reading it throws.
    bool live() const
    {
        if (!store_) return false;
        if (*alive_) return true;
        for (auto st : state_)
            if (st == DeviceNewer) throw std::runtime_error("wavefield read after its device state was destroyed");
        return false;
    }
    void alloc() const
    {
        if (data_.empty() && rows_ > 0 && cols_ > 0) data_.assign((size_t)(rows_ * cols_), fill_);
    }
    void pull_col(int c) const
    {
        std::vector<T> buf((size_t)rows_);
        store_->pull(c, buf.data());
        for (int64_t r = 0; r < rows_; ++r) data_[(size_t)(r * cols_ + c)] = buf[(size_t)r];
        state_[(size_t)c] = Synced;
    }
    void pull_stale() const
    {
        if (!live()) return;
        for (int c = 0; c < cols_; ++c)
            if (state_[(size_t)c] == DeviceNewer) {
                alloc();
                pull_col(c);
            }
    }
    void touch(int c, bool write) const
    {
        if (c < 0 || c >= cols_) throw std::out_of_range("Array2D column index");
        alloc();
        if (!live()) return;
        if (state_[(size_t)c] == DeviceNewer) pull_col(c);
        if (write) state_[(size_t)c] = HostNewer;
    }

    int64_t rows_ = 0;
    int cols_ = 0;
    T fill_ = T(0);
    mutable std::vector<T> data_;
    mutable std::vector<State> state_;
    ColumnStore<T> *store_ = nullptr;
    std::shared_ptr<bool> alive_;
};

using vectorReal = Vector<float>;
using vectorInt = Vector<int>;
using vectorDouble = Vector<double>;
using arrayInt = Array2D<int>;
using arrayReal = Array2D<float>;
using arrayDouble = Array2D<double>;

// dataType.hpp's allocation helpers (the name is the reference's debugging label)
template <class A> A allocateVector(int n, const char * = nullptr) { return A(n); }
template <class A> A allocateArray2D(int rows, int cols, const char * = nullptr) { return A(rows, cols); }
