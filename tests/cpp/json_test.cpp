// Host-only checks of the checkpoint metadata reader and writer (fluidsim_amd/csrc/fs_json.hpp):
// every finite double survives the text round trip bit for bit, the JSON grammar is enforced, and
// malformed text is rejected with std::runtime_error.  Prints one PASS line per group.
#include "fluidsim_amd/csrc/fs_json.hpp"

#include <cstdint>
#include <cstring>
#include <iostream>
#include <limits>
#include <random>
#include <string>

namespace {

int failures = 0;

void require(bool ok, const std::string& what)
{
    if (!ok) {
        std::cerr << "FAIL " << what << '\n';
        ++failures;
    }
}

bool rejects(const std::string& text)
{
    try {
        (void)fsjson::parse(text.data(), text.size());
    } catch (const std::runtime_error&) {
        return true;
    }
    return false;
}

std::uint64_t bits(double x)
{
    std::uint64_t b;
    std::memcpy(&b, &x, sizeof b);
    return b;
}

}  // namespace

int main()
{
    // 1. numbers: random bit patterns, extremes, subnormals, negative zero
    {
        std::mt19937_64 rng(12345);
        int tested = 0;
        std::vector<double> values = {0.0, -0.0, 1.0, -1.0, 0.1, 1e-300, 5e-324, 2.2250738585072014e-308,
            std::numeric_limits<double>::max(), -std::numeric_limits<double>::max(), 0.02, 1.0 / 3.0};
        for (int i = 0; i < 200000; ++i) {
            std::uint64_t b = rng();
            double x;
            std::memcpy(&x, &b, sizeof x);
            if (std::isfinite(x))
                values.push_back(x);
        }
        for (double x : values) {
            std::string text;
            fsjson::put_number(text, x);
            const fsjson::Value v = fsjson::parse(text.data(), text.size());
            require(v.kind == fsjson::Value::Number && bits(v.number) == bits(x), "round trip of " + text);
            ++tested;
        }
        std::string nan;
        fsjson::put_number(nan, std::numeric_limits<double>::quiet_NaN());
        require(nan == "null", "non-finite numbers are written as null");
        std::cout << "PASS " << tested << " doubles round-trip bit for bit\n";
    }

    // 2. structure: objects, arrays, strings with escapes, literals, member lookup
    {
        std::string text = "{";
        bool first = true;
        fsjson::put_key(text, "name", first);
        fsjson::put_string(text, "a\"b\\c\n\x01");
        fsjson::put_key(text, "list", first);
        text += "[1, 2.5e3, -0.0, true, false, null, {\"k\": []}]";
        fsjson::put_key(text, "n", first);
        fsjson::put_int(text, -42);
        text += "}";
        const fsjson::Value v = fsjson::parse(text.data(), text.size());
        require(v.kind == fsjson::Value::Object && v.members.size() == 3, "object with three members");
        require(v.find("name") && v.find("name")->text == "a\"b\\c\n\x01", "escaped string round trip");
        const fsjson::Value* list = v.find("list");
        require(list && list->items.size() == 7 && list->items[1].number == 2500.0 &&
                    list->items[3].kind == fsjson::Value::Bool && list->items[3].boolean &&
                    list->items[5].kind == fsjson::Value::Null && list->items[6].kind == fsjson::Value::Object,
            "array contents");
        require(v.find("n") && v.find("n")->number == -42.0, "integer member");
        require(v.find("missing") == nullptr, "absent member");
        const std::string u = "\"\\u0041\\u00e9\"";
        require(fsjson::parse(u.data(), u.size()).text == "A\xc3\xa9", "unicode escapes");
        std::cout << "PASS nested values, escapes and lookups\n";
    }

    // 3. grammar violations are rejected
    {
        const char* bad[] = {"", "{", "}", "[1,]", "{\"a\":1,}", "{\"a\" 1}", "{a:1}", "01", "1.", ".5", "-", "1e",
            "+1", "tru", "nul", "\"abc", "\"\\x\"", "[1 2]", "{\"a\":1}{", "{\"a\":1,\"a\":2}", "NaN", "Infinity",
            "\"\x01\""};
        int n = 0;
        for (const char* t : bad) {
            require(rejects(t), std::string("rejects ") + t);
            ++n;
        }
        std::string deep(100, '[');
        deep += std::string(100, ']');
        require(rejects(deep), "rejects nesting deeper than 32");
        std::cout << "PASS " << n + 1 << " malformed documents rejected\n";
    }

    return failures == 0 ? 0 : 1;
}
